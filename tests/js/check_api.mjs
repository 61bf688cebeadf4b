// Loads the JS binding and reports its surface: the WebGPUEngine method
// names the reference's callers use, and initGPU()'s result on this host.
import { WebGPUEngine, version, deviceCount } from '../../efs_amd/js/webgpu-engine.mjs';

const methods = ['initGPU', 'init', 'configureExcitation', 'configurePML', 'configureLorentzADE', 'configureTFSF',
    'configureRLC', 'configureMur', 'iterate', 'getFields', 'uploadFields', 'readProbe', 'configureProbeGather',
    'readProbeGather', 'configureProbeGatherBuffered', 'readProbeGatherBuffered', 'configureNF2FFAccumulation',
    'readNF2FFAccumulation', 'computeNF2FFfarField', 'computeEnergy', 'flushEnergy', 'destroy',
    'configurePBC', 'configureSteadyState', 'readSteadyState', 'checkConvergence'];
const eng = new WebGPUEngine();
const missing = methods.filter((m) => typeof eng[m] !== 'function');
eng.initGPU().then((ok) => {
    console.log(JSON.stringify({ version: version(), devices: deviceCount(), missing, initGPU: ok, numTS: eng.numTS }));
});
