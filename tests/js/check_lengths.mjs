// Calls the N-API addon with typed arrays of the wrong length for every
// read and upload that writes or reads a caller buffer, and reports which
// calls threw (all of them must).  Used by tests/test_napi.py.
import { createRequire } from 'module';
import { writeFileSync } from 'fs';

const require = createRequire(import.meta.url);
const addon = require('../../efs_amd/js/efs_engine.node');

function throws(fn) {
    try { fn(); } catch (e) { return String(e.message || e); }
    return null;
}

const h = addon.create(0);
if (!h) throw new Error('no GPU');
const n = 8, N3 = 3 * n * n * n;
const ones = new Float32Array(N3).fill(1);
const vi = new Float32Array(N3).fill(0.1);
addon.init(h, n, n, n, ones, vi, ones, vi);
addon.cfgGather(h, Uint32Array.from([1, 2, 3]));
addon.cfgGatherBuffered(h, Uint32Array.from([1, 2]), 16, 1);
const pts = Uint32Array.from([2, 2, 2, 3, 2, 2]);
const el = new Float32Array(6 * n).fill(1e-3);
addon.cfgNF2FF(h, pts, Float64Array.from([1e9]), 1e-12, el, Uint32Array.from([n, n, n, n, n, n]), 0, 0, 0);
addon.iterate(h, 4);
addon.sync(h);
const res = {
    getFieldsShort: throws(() => addon.getFields(h, new Float32Array(N3 - 1), null)),
    getFieldsLong: throws(() => addon.getFields(h, null, new Float32Array(N3 + 3))),
    uploadShort: throws(() => addon.uploadFields(h, new Float32Array(10), null)),
    readGatherShort: throws(() => addon.readGather(h, new Float32Array(5))),
    readGatherBufferedShort: throws(() => addon.readGatherBuffered(h, new Float32Array(4 * 4 - 1))),
    readNF2FFShort: throws(() => addon.readNF2FF(h, new Float32Array(11), new Float32Array(12))),
    farfieldMetaShort: throws(() => addon.farfield(h, 0, new Float32Array(15), Float32Array.from([0.5]),
        Float32Array.from([0.5]), Float64Array.from([0, 0, 0]), 1e9, 1, new Float32Array(5))),
};
// the right lengths still work
addon.getFields(h, new Float32Array(N3), new Float32Array(N3));
addon.readGather(h, new Float32Array(6));
const ns = addon.readGatherBuffered(h, new Float32Array(4 * 4));
addon.readNF2FF(h, new Float32Array(12), new Float32Array(12));
res.samples = ns;
addon.destroy(h);
writeFileSync(process.argv[2], JSON.stringify(res));
