// Runs computeNF2FF over surface fields through the JS module on the GPU
// and writes its result as JSON (tests/test_gpu_nf2ff_surface.py); with a
// third argument "points", writes only the surface points it builds, on the
// CPU (tests/test_nf2ff_surface.py).
import { readFileSync, writeFileSync } from 'fs';
import { computeNF2FF, surfacePoints } from '../../efs_amd/js/nf2ff.mjs';

const c = JSON.parse(readFileSync(process.argv[2], 'utf8'));
const faces = c.faces.map((f) => ({
    E: f.E.map((a) => Float64Array.from(a)),
    H: f.H.map((a) => Float64Array.from(a)),
    mesh: { x: Float64Array.from(f.mesh.x), y: Float64Array.from(f.mesh.y), z: Float64Array.from(f.mesh.z) },
    normal: f.normal,
    meshType: f.meshType,
}));
if (process.argv[4] === 'points') {
    const opts = c.opts || {};
    const p = surfacePoints({ faces }, opts.meshType || 0, opts.mirror || null);
    writeFileSync(process.argv[3], JSON.stringify({ meta: Array.from(p.meta), accE: Array.from(p.accE),
        accH: Array.from(p.accH), power: p.power }));
    process.exit(0);
}
const r = computeNF2FF({ faces }, c.freq, Float64Array.from(c.theta), Float64Array.from(c.phi), c.center, c.radius,
    c.opts);
const out = {};
for (const k of Object.keys(r)) out[k] = typeof r[k] === 'number' ? r[k] : Array.from(r[k]);
writeFileSync(process.argv[3], JSON.stringify(out));
