// batch-engine.mjs -- the batched form (include/efs_batch.h) for JavaScript.
//
// Many independent grids (instances) advanced together on one GPU.  Each
// instance is configured with the arguments a WebGPUEngine takes for the
// same call (init coefficients, configureMur, configureExcitation, the
// probe indices of configureProbeGatherBuffered); after finalize() every
// instance steps in lockstep, and its probe samples and fields read back in
// the WebGPUEngine shapes (readProbeGatherBuffered, getFields).
// Syntax is kept to what Node 12 parses.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const addon = require('./efs_engine.node');

const f32 = (a) => (a instanceof Float32Array ? a : Float32Array.from(a));
const u32 = (a) => (a instanceof Uint32Array ? a : Uint32Array.from(a));

export class BatchEngine {
    // options.storage: 'fp32' (default) or 'fp16'
    constructor(options) {
        this._opts = options || {};
        this._h = null;
        this._grids = [];
        this._nGather = [];
    }

    // false when no GPU is present (as WebGPUEngine.initGPU)
    async initGPU() {
        const storage = this._opts.storage === 'fp16' ? 1 : 0;
        const h = addon.batchCreate(this._opts.deviceIndex || 0, storage);
        if (!h) return false;
        this._h = h;
        return true;
    }

    _need() {
        if (!this._h) throw new Error('HIP device not initialized. Call initGPU() first.');
    }

    // init(gridSize, {vv, vi, ii, iv}) of a new instance; returns its id
    add(gridSize, coefficients) {
        this._need();
        const [nx, ny, nz] = gridSize;
        const id = addon.batchAdd(this._h, nx, ny, nz, f32(coefficients.vv), f32(coefficients.vi),
            f32(coefficients.ii), f32(coefficients.iv));
        this._grids[id] = [nx, ny, nz];
        this._nGather[id] = 0;
        return id;
    }

    configureExcitation(id, cfg) {
        this._need();
        addon.batchCfgExcitation(this._h, id, f32(cfg.signal), u32(cfg.delay), f32(cfg.amp), u32(cfg.dir),
            u32(cfg.pos), (cfg.period || 0) >>> 0);
    }

    // single list {coeff, normal_idx, shifted_idx} or the dual nyP / nyPP form
    configureMur(id, cfg) {
        this._need();
        let coeff, nidx, sidx;
        if (cfg.coeff_nyP) {
            const cat = (a, b, T) => { const o = new T(a.length + b.length); o.set(a, 0); o.set(b, a.length); return o; };
            coeff = cat(f32(cfg.coeff_nyP), f32(cfg.coeff_nyPP), Float32Array);
            nidx = cat(u32(cfg.normal_idx_nyP), u32(cfg.normal_idx_nyPP), Uint32Array);
            sidx = cat(u32(cfg.shifted_idx_nyP), u32(cfg.shifted_idx_nyPP), Uint32Array);
        } else {
            nidx = u32(cfg.normal_idx);
            sidx = u32(cfg.shifted_idx);
            coeff = typeof cfg.coeff === 'number' ? new Float32Array(nidx.length).fill(cfg.coeff) : f32(cfg.coeff);
        }
        addon.batchCfgMur(this._h, id, coeff, nidx, sidx);
    }

    configureProbeGather(id, indices) {
        this._need();
        const idx = u32(indices);
        addon.batchCfgGather(this._h, id, idx);
        this._nGather[id] = idx.length;
    }

    finalize(sampleInterval, maxSamples) {
        this._need();
        addon.batchFinalize(this._h, (sampleInterval || 1) >>> 0, (maxSamples || 0) >>> 0);
    }

    // autotune the block table and occupancy; leaves the batch reset
    tune(steps) {
        this._need();
        const r = addon.batchTune(this._h, (steps || 8) >>> 0);
        return { blocks: r[0], waves: r[1] };
    }

    reset() { this._need(); addon.batchReset(this._h); }

    iterate(n) { this._need(); addon.batchIterate(this._h, n >>> 0); }

    sync() { this._need(); addon.batchSync(this._h); return Promise.resolve(); }

    get numTS() { return this._h ? addon.batchNumTS(this._h) : 0; }

    info() {
        this._need();
        const r = addon.batchInfo(this._h);
        return { blocks: r[0], waves: r[1], graphSteps: r[2], cells: r[3], instances: this._grids.length };
    }

    // {data, numSamples, stride} of instance id
    async readProbeGatherBuffered(id) {
        this._need();
        const stride = 2 * this._nGather[id];
        const ns = addon.batchReadGather(this._h, id, null);
        const data = new Float32Array(Math.max(ns * stride, 1));
        if (ns && stride) addon.batchReadGather(this._h, id, data);
        return { data: data.subarray(0, ns * stride), numSamples: ns, stride };
    }

    async getFields(id) {
        this._need();
        const [nx, ny, nz] = this._grids[id];
        const volt = new Float32Array(3 * nx * ny * nz);
        const curr = new Float32Array(3 * nx * ny * nz);
        addon.batchGetFields(this._h, id, volt, curr);
        return { volt, curr };
    }

    destroy() {
        if (this._h) addon.batchDestroy(this._h);
        this._h = null;
    }
}
