// Writes tests/golden/reference_step_vectors.json: input/output vectors of
// the reference's CPUFDTDEngine.step() composition (src/webgpu-fdtd.mjs:
// 1094-1172) on small seeded grids with every CPU-engine extension active at
// once (UPML, excitation, Mur, Lorentz/Drude ADE, TFSF, lumped RLC).  The
// reference's own tests pin each extension alone; nothing in them pins the
// composed multi-step result, so these vectors are generated by running the
// reference function itself.  Run (Node >= 12) from the repository root:
//     node tests/golden/gen_reference_vectors.mjs [path/to/reference/src]
// The output holds only numbers (inputs and expected outputs).
import { writeFileSync } from 'fs';
import { fileURLToPath, pathToFileURL } from 'url';
import { dirname, join } from 'path';

const here = dirname(fileURLToPath(import.meta.url));
const refSrc = process.argv[2] || '/root/reference/src';

function mulberry32(a) {
    return function () {
        a |= 0; a = (a + 0x6D2B79F5) | 0;
        let t = Math.imul(a ^ (a >>> 15), 1 | a);
        t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t;
        return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
    };
}

function f32(n, rnd, lo, hi) {
    const a = new Float32Array(n);
    for (let i = 0; i < n; i++) a[i] = lo + (hi - lo) * rnd();
    return a;
}

const arr = (a) => Array.from(a);

function coeffs(N, rnd) {
    const t = 3 * N;
    return { vv: f32(t, rnd, 0.9, 1.0), vi: f32(t, rnd, 0.05, 0.3), ii: f32(t, rnd, 0.9, 1.0), iv: f32(t, rnd, 0.05, 0.3) };
}

function regionCoeffs(n, rnd) {
    return {
        vv: f32(n, rnd, 0.6, 1.0), vvfo: f32(n, rnd, 0.0, 0.3), vvfn: f32(n, rnd, 0.8, 1.4),
        ii: f32(n, rnd, 0.6, 1.0), iifo: f32(n, rnd, 0.0, 0.3), iifn: f32(n, rnd, 0.8, 1.4),
    };
}

async function main() {
    const { CPUFDTDEngine } = await import(pathToFileURL(join(refSrc, 'webgpu-fdtd.mjs')).href);
    const cases = [];

    // --- case "core": random coefficients and fields, core updates only ---
    {
        const rnd = mulberry32(20261203);
        const dims = [5, 4, 6];
        const N = dims[0] * dims[1] * dims[2];
        const c = coeffs(N, rnd);
        const volt0 = f32(3 * N, rnd, -1, 1), curr0 = f32(3 * N, rnd, -1, 1);
        const eng = new CPUFDTDEngine(dims, c);
        eng.uploadFields(volt0, curr0);
        eng.iterate(6);
        cases.push({ name: 'core', dims, steps: 6, coeffs: { vv: arr(c.vv), vi: arr(c.vi), ii: arr(c.ii), iv: arr(c.iv) },
            volt0: arr(volt0), curr0: arr(curr0), out: { volt: arr(eng.volt), curr: arr(eng.curr), numTS: eng.numTS } });
    }

    // --- case "all": every extension at once ---
    for (const [name, seed, dims, steps, period] of [['all', 7, [7, 6, 5], 15, 0], ['all_periodic', 11, [6, 7, 6], 21, 5]]) {
        const rnd = mulberry32(seed);
        const [Nx, Ny, Nz] = dims;
        const N = Nx * Ny * Nz;
        const lin = (x, y, z) => x * Ny * Nz + y * Nz + z;
        const gidx = (n, x, y, z) => n * N + lin(x, y, z);
        const c = coeffs(N, rnd);
        const volt0 = f32(3 * N, rnd, -0.5, 0.5), curr0 = f32(3 * N, rnd, -0.5, 0.5);

        // Two non-overlapping UPML regions: x-min slab and a z-max slab.
        const r0 = { startPos: [0, 0, 0], numLines: [2, Ny, Nz] };
        const r1 = { startPos: [2, 0, Nz - 2], numLines: [Nx - 2, Ny, 2] };
        for (const r of [r0, r1]) Object.assign(r, regionCoeffs(3 * r.numLines[0] * r.numLines[1] * r.numLines[2], rnd));

        const signal = f32(9, rnd, -1, 1);
        const exc = { signal, delay: new Uint32Array([0, 2, 30]), amp: new Float32Array([1.5, -0.7, 0.9]),
            dir: new Uint32Array([2, 0, 1]), pos: new Uint32Array([lin(3, 3, 2), lin(4, 2, 1), lin(1, 4, 3)]), period };

        // Mur on the y = Ny-1 face (normal y = Ny-1, shifted y = Ny-2), tangential comps 0 and 2.
        const murN = [], murS = [], murC = [];
        for (let x = 0; x < Nx; x += 2) for (let z = 0; z < Nz; z += 2) for (const n of [0, 2]) {
            murN.push(gidx(n, x, Ny - 1, z)); murS.push(gidx(n, x, Ny - 2, z)); murC.push(0.2 + 0.3 * rnd());
        }
        const mur = { coeff: new Float32Array(murC), normal_idx: new Uint32Array(murN), shifted_idx: new Uint32Array(murS) };

        // ADE: a Drude order and a Lorentz order; one cell lies in UPML region 0, one on a Mur face.
        const adePos = new Uint32Array([lin(1, 2, 2), lin(3, 2, 2), lin(4, 3, 1), lin(2, Ny - 1, 2)]);
        const mkDir = (dir) => ({ dir, pos_idx: adePos,
            v_int_ADE: f32(4, rnd, 0.5, 0.95), v_ext_ADE: f32(4, rnd, 0.05, 0.3), v_Lor_ADE: f32(4, rnd, 0.0, 0.2),
            i_int_ADE: f32(4, rnd, 0.5, 0.95), i_ext_ADE: f32(4, rnd, 0.05, 0.3), i_Lor_ADE: f32(4, rnd, 0.0, 0.2) });
        const ade = { orders: [
            { numCells: 4, hasLorentz: false, directions: [mkDir(0), mkDir(2)] },
            { numCells: 4, hasLorentz: true, directions: [mkDir(1)] },
        ] };

        const tsig = f32(12, rnd, -1, 1);
        const tpt = (fi, di) => ({ field_idx: fi, delay_int: di, delay_frac: Math.fround(rnd()), amp: Math.fround(0.5 + rnd()) });
        const tfsf = { signal: tsig, period,
            voltagePoints: [tpt(gidx(1, 3, 3, 2), 0), tpt(gidx(2, 4, 2, 2), 3), tpt(gidx(0, 2, 3, 1), 1)],
            currentPoints: [tpt(gidx(0, 3, 2, 2), 1), tpt(gidx(2, 5, 3, 1), 2)] };

        const rl = () => Math.fround(rnd());
        const rlc = { elements: [
            { field_idx: lin(3, 1, 3), direction: 2, type_flag: 0, i2v: rl(), ilv: rl(), vvd: 0, vv2: 0, vj1: 0, vj2: 0, ib0: 0, b1: 0, b2: 0 },
            { field_idx: lin(4, 4, 2), direction: 1, type_flag: 1, i2v: 0, ilv: 0, vvd: 0.5 + 0.4 * rl(), vv2: 0.2 * rl(),
              vj1: 0.1 * rl(), vj2: 0.1 * rl(), ib0: 0.3 * rl(), b1: 0.2 * rl(), b2: 0.1 * rl() },
        ] };

        const eng = new CPUFDTDEngine(dims, c);
        eng.uploadFields(volt0, curr0);
        eng.configurePML([r0, r1]);
        eng.configureExcitation(exc);
        eng.configureMur(mur);
        eng.configureLorentz(ade);
        eng.configureTFSF(tfsf);
        eng.configureRLC(rlc);
        eng.iterate(steps);

        const ser = (o) => { const r = {}; for (const k of Object.keys(o)) r[k] = ArrayBuffer.isView(o[k]) ? arr(o[k]) : o[k]; return r; };
        cases.push({
            name, dims, steps,
            coeffs: { vv: arr(c.vv), vi: arr(c.vi), ii: arr(c.ii), iv: arr(c.iv) },
            volt0: arr(volt0), curr0: arr(curr0),
            pml: [ser(r0), ser(r1)], excitation: ser(exc), mur: ser(mur),
            ade: { orders: ade.orders.map((o) => ({ numCells: o.numCells, hasLorentz: o.hasLorentz, directions: o.directions.map(ser) })) },
            tfsf: { signal: arr(tsig), period, voltagePoints: tfsf.voltagePoints, currentPoints: tfsf.currentPoints },
            rlc,
            out: {
                volt: arr(eng.volt), curr: arr(eng.curr), numTS: eng.numTS,
                pml_flux: eng.pmlRegions.map((r) => ({ volt: arr(r.volt_flux), curr: arr(r.curr_flux) })),
                ade: eng.lorentzOrders.map((o) => o.directions.map((d) => ({
                    volt_ADE: arr(d.volt_ADE), volt_Lor_ADE: arr(d.volt_Lor_ADE), curr_ADE: arr(d.curr_ADE), curr_Lor_ADE: arr(d.curr_Lor_ADE) }))),
                rlc: { Vdn: arr(eng.rlcVdn), Jn: arr(eng.rlcJn), VIl: arr(eng.rlcVIl) },
            },
        });
    }

    const out = { generator: 'tests/golden/gen_reference_vectors.mjs', source: 'CPUFDTDEngine (src/webgpu-fdtd.mjs)', cases };
    writeFileSync(join(here, 'reference_step_vectors.json'), JSON.stringify(out));
    console.log(`wrote ${cases.length} cases`);
}

main().catch((e) => { console.error(e); process.exit(1); });
