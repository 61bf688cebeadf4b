// Runs one case through the JS binding on the GPU: reads a JSON case
// (dims, coeffs, volt0/curr0, pml, excitation, mur, steps), writes the
// resulting fields and energy as JSON.  Used by tests/test_napi.py.
import { readFileSync, writeFileSync } from 'fs';
import { WebGPUEngine } from '../../efs_amd/js/webgpu-engine.mjs';

async function main() {
    const c = JSON.parse(readFileSync(process.argv[2], 'utf8'));
    const eng = new WebGPUEngine();
    if (!(await eng.initGPU())) throw new Error('no GPU');
    await eng.init(c.dims, c.coeffs);
    eng.uploadFields(c.volt0, c.curr0);
    if (c.pml) eng.configurePML(c.pml);
    if (c.excitation) eng.configureExcitation(c.excitation);
    if (c.mur) eng.configureMur(c.mur);
    if (c.ade) eng.configureLorentzADE(c.ade);
    if (c.tfsf) eng.configureTFSF(c.tfsf);
    if (c.rlc) eng.configureRLC(c.rlc);
    eng.iterate(c.steps);
    await eng.device.queue.onSubmittedWorkDone();
    const f = await eng.getFields();
    const out = { numTS: eng.numTS, volt: Array.from(f.volt), curr: Array.from(f.curr), energy: eng.energy() };
    eng.destroy();
    writeFileSync(process.argv[3], JSON.stringify(out));
}
main().catch((e) => { console.error(e); process.exit(1); });
