// Loads the JS binding and reports its surface: the WebGPUEngine method
// names the reference's callers use, and initGPU()'s result on this host.
import { WebGPUEngine, version, deviceCount } from '../../efs_amd/js/webgpu-engine.mjs';
import { BatchEngine } from '../../efs_amd/js/batch-engine.mjs';

const methods = ['initGPU', 'init', 'configureExcitation', 'configurePML', 'configureLorentzADE', 'configureTFSF',
    'configureRLC', 'configureMur', 'iterate', 'getFields', 'uploadFields', 'readProbe', 'configureProbeGather',
    'readProbeGather', 'configureProbeGatherBuffered', 'readProbeGatherBuffered', 'configureNF2FFAccumulation',
    'readNF2FFAccumulation', 'computeNF2FFfarField', 'computeEnergy', 'flushEnergy', 'destroy',
    'configurePBC', 'configureSteadyState', 'readSteadyState', 'checkConvergence'];
const batchMethods = ['initGPU', 'add', 'configureExcitation', 'configureMur', 'configureProbeGather', 'finalize',
    'tune', 'reset', 'iterate', 'sync', 'info', 'readProbeGatherBuffered', 'getFields', 'destroy'];
const eng = new WebGPUEngine();
const bat = new BatchEngine();
const missing = methods.filter((m) => typeof eng[m] !== 'function')
    .concat(batchMethods.filter((m) => typeof bat[m] !== 'function').map((m) => 'BatchEngine.' + m));
Promise.all([eng.initGPU(), bat.initGPU()]).then(([ok, okb]) => {
    console.log(JSON.stringify({ version: version(), devices: deviceCount(), missing, initGPU: ok, batchInitGPU: okb,
        numTS: eng.numTS }));
});
