// Runs a batch of instances through the JS BatchEngine on the GPU: reads a
// JSON list of instances (dims, coeffs, mur, excitation, probes) plus
// steps, writes each instance's fields and gather samples as JSON.  Used by
// tests/test_napi.py.
import { readFileSync, writeFileSync } from 'fs';
import { BatchEngine } from '../../efs_amd/js/batch-engine.mjs';

async function main() {
    const c = JSON.parse(readFileSync(process.argv[2], 'utf8'));
    const b = new BatchEngine({ storage: c.storage });
    if (!(await b.initGPU())) throw new Error('no GPU');
    for (const inst of c.instances) {
        const id = b.add(inst.dims, inst.coeffs);
        b.configureMur(id, inst.mur);
        b.configureExcitation(id, inst.excitation);
        b.configureProbeGather(id, inst.probes);
    }
    b.finalize(1, c.steps);
    b.iterate(c.steps);
    await b.sync();
    const out = { numTS: b.numTS, info: b.info(), instances: [] };
    for (let id = 0; id < c.instances.length; ++id) {
        const f = await b.getFields(id);
        const g = await b.readProbeGatherBuffered(id);
        out.instances.push({ volt: Array.from(f.volt), curr: Array.from(f.curr), gather: Array.from(g.data),
            numSamples: g.numSamples });
    }
    b.destroy();
    writeFileSync(process.argv[3], JSON.stringify(out));
}
main().catch((e) => { console.error(e); process.exit(1); });
