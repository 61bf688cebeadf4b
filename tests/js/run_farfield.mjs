// Runs a small NF2FF case through the JS binding on the GPU and writes the
// far field with and without the mirror option as JSON.  Used by
// tests/test_napi.py.
import { readFileSync, writeFileSync } from 'fs';
import { WebGPUEngine } from '../../efs_amd/js/webgpu-engine.mjs';

async function main() {
    const c = JSON.parse(readFileSync(process.argv[2], 'utf8'));
    const eng = new WebGPUEngine();
    if (!(await eng.initGPU())) throw new Error('no GPU');
    await eng.init(c.dims, c.coeffs);
    eng.configureExcitation(c.excitation);
    eng.configureNF2FFAccumulation(c.nf2ff);
    eng.iterate(c.steps);
    await eng.device.queue.onSubmittedWorkDone();
    const cfg = { pointMeta: c.pointMeta, theta: c.theta, phi: c.phi, center: c.center, frequency: c.frequency,
        radius: 1.0 };
    const plain = await eng.computeNF2FFfarField(cfg);
    const mirrored = await eng.computeNF2FFfarField(Object.assign({}, cfg, { mirror: c.mirror }));
    eng.destroy();
    writeFileSync(process.argv[3], JSON.stringify({ plain: Array.from(plain), mirrored: Array.from(mirrored) }));
}
main().catch((e) => { console.error(e); process.exit(1); });
