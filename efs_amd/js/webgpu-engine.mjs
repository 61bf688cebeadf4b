// webgpu-engine.mjs -- WebGPUEngine-compatible class over the MI355X engine.
//
// Drop-in for the reference's `WebGPUEngine` (src/webgpu-engine.mjs:994-3498):
// same method names, argument objects and return shapes; reads return
// Promises.  Calls go through the N-API addon (efs_engine.node) to the C ABI
// in include/efs_engine.h.  Syntax is kept to what Node 12 parses.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const addon = require('./efs_engine.node');

const f32 = (a) => (a instanceof Float32Array ? a : Float32Array.from(a));
const u32 = (a) => (a instanceof Uint32Array ? a : Uint32Array.from(a));
const f64 = (a) => (a instanceof Float64Array ? a : Float64Array.from(a));
const RLC_KEYS = ['i2v', 'ilv', 'vvd', 'vv2', 'vj1', 'vj2', 'ib0', 'b1', 'b2'];

function concatF32(list) {
    let n = 0;
    for (const a of list) n += a.length;
    const out = new Float32Array(n);
    let o = 0;
    for (const a of list) { out.set(a, o); o += a.length; }
    return out;
}

export class HIPEngine {
    constructor(deviceIndex) {
        this.deviceIndex = deviceIndex || 0;
        this.device = null;
        this.numLines = null;
        this.totalCells = 0;
        this._h = null;
        this._gatherCount = 0;
        this._gatherBuffered = null;
        this._nf2ff = null;
        this._energyPending = null;
        this._energyLastValue = 0;
    }

    // initGPU() (src/webgpu-engine.mjs:1107-1135): false when no GPU.
    async initGPU() {
        const h = addon.create(this.deviceIndex);
        if (!h) return false;
        this._h = h;
        const self = this;
        this.device = { queue: { onSubmittedWorkDone() { addon.sync(self._h); return Promise.resolve(); } } };
        return true;
    }

    _need() {
        if (!this._h) throw new Error('HIP device not initialized. Call initGPU() first.');
    }

    get numTS() { return this._h ? addon.numTS(this._h) : 0; }
    set numTS(v) { this._need(); addon.setNumTS(this._h, v >>> 0); }

    // options.storage: 'fp32' (default), 'fp16' (fields stored as FP16
    // with per-component power-of-two scales) or 'fp64' (the FP64 run of the
    // same engine); reads return FP32.
    async init(gridSize, coefficients, options) {
        this._need();
        const storage = (options && options.storage) || 'fp32';
        const codes = { fp32: 0, fp16: 1, fp64: 2 };
        if (!(storage in codes)) throw new Error("storage must be 'fp32', 'fp16' or 'fp64'");
        addon.setStorage(this._h, codes[storage]);
        const [Nx, Ny, Nz] = gridSize;
        addon.init(this._h, Nx, Ny, Nz, f32(coefficients.vv), f32(coefficients.vi), f32(coefficients.ii), f32(coefficients.iv));
        this.numLines = [Nx, Ny, Nz];
        this.totalCells = Nx * Ny * Nz;
        this._gatherCount = 0;
        this._gatherBuffered = null;
        this._nf2ff = null;
        this._energyPending = null;
        this._energyLastValue = 0;
    }

    configureExcitation(e) {
        this._need();
        addon.cfgExcitation(this._h, f32(e.signal), u32(e.delay), f32(e.amp), u32(e.dir), u32(e.pos), (e.period || 0) >>> 0);
    }

    configurePML(regions) {
        this._need();
        const start = [], num = [];
        const cat = { vv: [], vvfo: [], vvfn: [], ii: [], iifo: [], iifn: [] };
        for (const r of regions) {
            start.push(...r.startPos);
            num.push(...r.numLines);
            for (const k of Object.keys(cat)) cat[k].push(f32(r[k]));
        }
        addon.cfgPML(this._h, u32(start), u32(num), concatF32(cat.vv), concatF32(cat.vvfo), concatF32(cat.vvfn),
            concatF32(cat.ii), concatF32(cat.iifo), concatF32(cat.iifn));
    }

    // configureLorentzADE(config) (src/webgpu-engine.mjs:1358-1475); a second
    // argument {apply: true} selects the openEMS write-back of the ADE state.
    configureLorentzADE(config, options) {
        this._need();
        addon.adeClear(this._h);
        for (const order of config.orders) {
            for (const d of order.directions) {
                const opt = (a) => (a ? f32(a) : null);
                addon.adeAdd(this._h, order.hasLorentz ? 1 : 0, d.dir, u32(d.pos_idx), opt(d.v_int_ADE), opt(d.v_ext_ADE),
                    opt(d.v_Lor_ADE), opt(d.i_int_ADE), opt(d.i_ext_ADE), opt(d.i_Lor_ADE));
            }
        }
        addon.adeApply(this._h, options && options.apply ? 1 : 0);
    }

    configureTFSF(config) {
        this._need();
        const cols = (pts) => [u32(pts.map((p) => p.field_idx)), u32(pts.map((p) => p.delay_int)),
            f32(pts.map((p) => p.delay_frac)), f32(pts.map((p) => p.amp))];
        const v = cols(config.voltagePoints || []);
        const c = cols(config.currentPoints || []);
        addon.cfgTFSF(this._h, f32(config.signal), (config.period || 0) >>> 0, v[0], v[1], v[2], v[3], c[0], c[1], c[2], c[3]);
    }

    configureRLC(config) {
        this._need();
        const el = config.elements;
        if (el.length === 0) return;
        const co = new Float32Array(9 * el.length);
        el.forEach((e, i) => RLC_KEYS.forEach((k, j) => { co[9 * i + j] = e[k] || 0; }));
        addon.cfgRLC(this._h, u32(el.map((e) => e.field_idx)), u32(el.map((e) => e.direction)),
            u32(el.map((e) => e.type_flag)), co);
    }

    // configureMur (src/webgpu-engine.mjs:1822-1918): single or dual form.
    configureMur(config) {
        this._need();
        let coeff, nidx, sidx;
        if (config.coeff_nyP) {
            coeff = concatF32([f32(config.coeff_nyP), f32(config.coeff_nyPP)]);
            nidx = u32([...config.normal_idx_nyP, ...config.normal_idx_nyPP]);
            sidx = u32([...config.shifted_idx_nyP, ...config.shifted_idx_nyPP]);
        } else {
            nidx = u32(config.normal_idx);
            sidx = u32(config.shifted_idx);
            coeff = typeof config.coeff === 'number' ? new Float32Array(nidx.length).fill(config.coeff) : f32(config.coeff);
        }
        addon.cfgMur(this._h, coeff, nidx, sidx);
    }

    // configurePBC({axes: [{axis, phase}]}) (src/webgpu-engine.mjs:1976-2155).
    configurePBC(config) {
        this._need();
        const axes = (config && config.axes) || [];
        if (axes.length === 0) return;
        addon.cfgPBC(this._h, u32(axes.map((a) => a.axis)), f64(axes.map((a) => a.phase || 0)));
    }

    // configureSteadyState({probe_idx, periodSamples, threshold})
    // (src/webgpu-engine.mjs:2161-2217); recording starts off.
    configureSteadyState(config) {
        this._need();
        const idx = u32(config.probe_idx);
        addon.cfgSteadyState(this._h, idx, config.periodSamples >>> 0);
        this._ss = { n: idx.length, period: config.periodSamples >>> 0, threshold: config.threshold || 1e-6 };
    }

    get ssRecording() { return this._h && this._ss ? addon.ssRead(this._h, null, null)[1] === 1 : false; }
    set ssRecording(on) { this._need(); addon.ssSet(this._h, on ? 1 : 0, addon.ssRead(this._h, null, null)[0]); }
    get ssCurrentSample() { return this._h && this._ss ? addon.ssRead(this._h, null, null)[0] : 0; }
    set ssCurrentSample(v) { this._need(); addon.ssSet(this._h, addon.ssRead(this._h, null, null)[1], v >>> 0); }

    fieldScales() {
        this._need();
        const out = new Float32Array(6);
        addon.fieldScales(this._h, out);
        return out;
    }

    // Period energies and counter (host copy).
    readSteadyState() {
        if (!this._ss) return null;
        const e1 = new Float32Array(this._ss.n), e2 = new Float32Array(this._ss.n);
        const [cur] = addon.ssRead(this._h, e1, e2);
        return { energy_period1: e1, energy_period2: e2, currentSample: cur };
    }

    // CPU engine's convergence rule (src/webgpu-fdtd.mjs:1077-1095).
    checkConvergence() {
        const s = this.readSteadyState();
        if (!s || s.currentSample < 2 * this._ss.period) return false;
        let maxRatio = 0;
        for (let n = 0; n < this._ss.n; n++) {
            const e1 = s.energy_period1[n], e2 = s.energy_period2[n];
            if (Math.abs(e1) < 1e-30) continue;
            const r = Math.abs(e2 - e1) / Math.abs(e1);
            if (r > maxRatio) maxRatio = r;
        }
        return maxRatio < this._ss.threshold;
    }

    iterate(numSteps) {
        this._need();
        addon.iterate(this._h, numSteps >>> 0);
    }

    // Not in the reference: picks the sweeps' workgroup target by timing
    // (efs_tune_sweeps) and leaves the fields as they were; returns the
    // chosen factor on the planner's target.
    tuneSweeps(reps) {
        this._need();
        return addon.tuneSweeps(this._h, reps === undefined ? 5 : reps >>> 0);
    }

    async getFields() {
        this._need();
        const volt = new Float32Array(3 * this.totalCells);
        const curr = new Float32Array(3 * this.totalCells);
        addon.getFields(this._h, volt, curr);
        return { volt, curr };
    }

    uploadFields(volt, curr) {
        this._need();
        addon.uploadFields(this._h, volt ? f32(volt) : null, curr ? f32(curr) : null);
    }

    async readProbe(positions) {
        const { volt, curr } = await this.getFields();
        const [Nx, Ny, Nz] = this.numLines;
        const out = new Float32Array(positions.length * 2);
        positions.forEach((p, i) => {
            const j = p.component * Nx * Ny * Nz + p.x * Ny * Nz + p.y * Nz + p.z;
            out[2 * i] = volt[j];
            out[2 * i + 1] = curr[j];
        });
        return out;
    }

    configureProbeGather(indices) {
        if (indices.length === 0) return;
        addon.cfgGather(this._h, u32(indices));
        this._gatherCount = indices.length;
    }

    async readProbeGather() {
        if (!this._gatherCount) return null;
        const out = new Float32Array(2 * this._gatherCount);
        addon.readGather(this._h, out);
        return out;
    }

    configureProbeGatherBuffered(indices, { maxSamples, sampleInterval }) {
        if (indices.length === 0) return;
        addon.cfgGatherBuffered(this._h, u32(indices), maxSamples >>> 0, sampleInterval >>> 0);
        this._gatherBuffered = { count: indices.length, maxSamples };
    }

    async readProbeGatherBuffered() {
        if (!this._gatherBuffered) return null;
        const ns = addon.readGatherBuffered(this._h, null);
        if (ns === 0) return null;
        const stride = 2 * this._gatherBuffered.count;
        const data = new Float32Array(ns * stride);
        addon.readGatherBuffered(this._h, data);
        return { data, numSamples: ns, stride };
    }

    // configureNF2FFAccumulation (src/webgpu-engine.mjs:2729-2827); omega may
    // be a number (reference form) or an array of angular frequencies.
    configureNF2FFAccumulation(config) {
        // the point count is the surface list's; a numPoints that disagrees
        // with it is refused rather than trusted for the read sizes
        const numPoints = Math.floor(config.surfaceIndices.length / 3);
        if (config.numPoints !== undefined && config.numPoints !== numPoints)
            throw new Error(`configureNF2FFAccumulation: numPoints ${config.numPoints} does not match ` +
                `surfaceIndices (${numPoints} points)`);
        if (numPoints === 0) return;
        const els = [...config.primalEdgeLens, ...config.dualEdgeLens].map(f32);
        const omegas = typeof config.omega === 'number' ? [config.omega] : Array.from(config.omega);
        addon.cfgNF2FF(this._h, u32(config.surfaceIndices), f64(omegas), config.dT, concatF32(els),
            u32(els.map((a) => a.length)), (config.maxTS || 0) >>> 0, (config.windowType || 0) >>> 0, (config.kBlock || 0) >>> 0,
            config.compensated === false ? 0 : 1);
        this._nf2ff = { numPoints, F: omegas.length, scalar: typeof config.omega === 'number' };
    }

    async readNF2FFAccumulation() {
        if (!this._nf2ff) return null;
        const { numPoints, F, scalar } = this._nf2ff;
        const accumE = new Float32Array(F * 6 * numPoints);
        const accumH = new Float32Array(F * 6 * numPoints);
        addon.readNF2FF(this._h, accumE, accumH);
        if (scalar) return { accumE, accumH };
        const split = (a) => Array.from({ length: F }, (_, f) => a.subarray(f * 6 * numPoints, (f + 1) * 6 * numPoints));
        return { accumE: split(accumE), accumH: split(accumH) };
    }

    // Optional config.mirror = { type: 'PEC' | 'PMC', direction, position,
    // imageNormals }: the surface's image in that plane (the mirror option of
    // src/nf2ff.mjs:377-398: the reflected faces keep their normals, as there;
    // imageNormals: true reverses the normals of reflected faces perpendicular
    // to the plane, as image theory has it).
    async computeNF2FFfarField(config) {
        const out = new Float32Array(5 * config.theta.length * config.phi.length);
        const m = config.mirror;
        if (!m) {
            addon.farfield(this._h, (config.freqIndex || 0) >>> 0, f32(config.pointMeta), f32(config.theta),
                f32(config.phi), f64(config.center), config.frequency, config.radius, out);
        } else {
            addon.farfield(this._h, (config.freqIndex || 0) >>> 0, f32(config.pointMeta), f32(config.theta),
                f32(config.phi), f64(config.center), config.frequency, config.radius, out,
                m.direction, m.position || 0, String(m.type || 'PEC').toUpperCase() === 'PMC' ? 1 : 0,
                m.imageNormals ? 1 : 0);
        }
        return out;
    }

    // computeEnergy() returns the previous call's value (0 first), as the
    // reference's double-buffered readback does (src/webgpu-engine.mjs:2974-3021).
    async computeEnergy() {
        const prev = this._energyPending;
        this._energyPending = addon.energy(this._h);
        if (prev !== null) this._energyLastValue = prev;
        return this._energyLastValue;
    }

    async flushEnergy() {
        if (this._energyPending !== null) {
            this._energyLastValue = this._energyPending;
            this._energyPending = null;
        }
        return this._energyLastValue;
    }

    // Synchronous current energy (FP64 reduction).
    energy() { return addon.energy(this._h); }

    destroy() {
        if (this._h) addon.destroy(this._h);
        this._h = null;
        this.device = null;
    }
}

export { HIPEngine as WebGPUEngine };
export const version = addon.version;
export const deviceCount = addon.deviceCount;
