// nf2ff.mjs -- computeNF2FF over surface fields, on the GPU.
//
// Drop-in for the reference's CPU far field (src/nf2ff.mjs:102-436):
// computeNF2FF(surfaceData, freq, theta, phi, center, radius, opts) with the
// same faces ({E, H, mesh, normal, meshType}), options (meshType, mirror)
// and result ({E_theta_re, E_theta_im, E_phi_re, E_phi_im, Prad, Dmax,
// P_rad}).  The faces become surface points here; the radiation integrals
// run on the GPU (efs_farfield_surface through the addon).
//
// A point of a cylindrical mesh carries (rho, alpha, z) components and its
// face normal in that frame; the reference rotates Js = n x H and Ms = -n x E
// by the point's alpha.  Rotation about z commutes with the cross product,
// so that is n' x H' with the rotated normal and fields, and since the
// integrals are linear in area * n, the point goes to the GPU as the
// Cartesian points of the components of n'.  Mirror passes
// (src/nf2ff.mjs:377-398) are added as reflected points with the PEC/PMC
// signs on the stored components and the normals kept; the radiated power
// is that of the unreflected surface.  Syntax is kept to what Node 12
// parses.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const addon = require('./efs_engine.node');

const C0 = 299792458;

function edgeLengths(lines) {
    const n = lines.length;
    const out = new Float64Array(n);
    if (n === 1) return out;
    out[0] = 0.5 * Math.abs(lines[1] - lines[0]);
    for (let i = 1; i < n - 1; i++) out[i] = 0.5 * Math.abs(lines[i + 1] - lines[i - 1]);
    out[n - 1] = 0.5 * Math.abs(lines[n - 1] - lines[n - 2]);
    return out;
}

// Surface points of all faces for one pass; push into the lists, return the
// pass's Poynting power.
function facePoints(faces, defaultMeshType, eSign, hSign, mirrorDir, mirrorPos, lists) {
    let power = 0;
    for (const face of faces) {
        const { E, H, mesh, normal } = face;
        const meshType = face.meshType !== undefined ? face.meshType : defaultMeshType;
        let ny = -1;
        for (let n = 0; n < 3; n++) if (normal[n] !== 0) { ny = n; break; }
        if (ny < 0) continue;
        const nP = (ny + 1) % 3, nPP = (ny + 2) % 3;
        const m = [mesh.x, mesh.y, mesh.z];
        const numP = m[nP].length, numPP = m[nPP].length;
        const sign = normal[ny];
        const eP = edgeLengths(m[nP]), ePP = edgeLengths(m[nPP]);
        const cyl = meshType === 1;
        const face_pts = [];  // per point: pos, rotated normal, area, E re/im, H re/im
        if (cyl && ny === 0) for (let i = 0; i < eP.length; i++) eP[i] *= m[0][0];
        for (let iP = 0; iP < numP; iP++) {
            for (let iPP = 0; iPP < numPP; iPP++) {
                const idx = ny === 1 ? iPP * numP + iP : iP * numPP + iPP;
                let area = eP[iP] * ePP[iPP];
                if (cyl && ny === 2) area *= m[0][iP];
                const er = [0, 0, 0], ei = [0, 0, 0], hr = [0, 0, 0], hi = [0, 0, 0];
                for (let c = 0; c < 3; c++) {
                    er[c] = E[c][2 * idx] * eSign[c]; ei[c] = E[c][2 * idx + 1] * eSign[c];
                    hr[c] = H[c][2 * idx] * hSign[c]; hi[c] = H[c][2 * idx + 1] * hSign[c];
                }
                if (mirrorDir < 0) {
                    const s = (er[nP] * hr[nPP] + ei[nP] * hi[nPP]) - (er[nPP] * hr[nP] + ei[nPP] * hi[nP]);
                    power += 0.5 * area * s * sign;
                }
                let pos = [0, 0, 0];
                pos[ny] = m[ny][0];
                pos[nP] = m[nP][iP];
                pos[nPP] = m[nPP][iPP];
                const nv = [0, 0, 0];
                nv[ny] = sign;
                if (cyl) {
                    const a = pos[1], ca = Math.cos(a), sa = Math.sin(a);
                    const rot = (v) => { const x = v[0] * ca - v[1] * sa, y = v[0] * sa + v[1] * ca; v[0] = x; v[1] = y; };
                    rot(er); rot(ei); rot(hr); rot(hi); rot(nv);
                    pos = [pos[0] * ca, pos[0] * sa, pos[2]];
                }
                if (mirrorDir >= 0) pos[mirrorDir] = 2 * mirrorPos - pos[mirrorDir];
                face_pts.push({ pos, nv, area, er, ei, hr, hi });
            }
        }
        // a rotated normal splits into its Cartesian components (axis-major
        // per face, the order of the Python binding)
        for (let ax = 0; ax < 3; ax++) {
            for (const q of face_pts) {
                if (q.nv[ax] === 0) continue;
                lists.meta.push(q.pos[0], q.pos[1], q.pos[2], ax, q.nv[ax] > 0 ? 1 : -1, q.area * Math.abs(q.nv[ax]),
                    0, 0);
                for (let c = 0; c < 3; c++) {
                    lists.e.push(q.er[c], q.ei[c]);
                    lists.h.push(q.hr[c], q.hi[c]);
                }
            }
        }
    }
    return power;
}

// The GPU point list of a surface: {meta, accE, accH, power}.
export function surfacePoints(surfaceData, defaultMeshType, mirror) {
    const lists = { meta: [], e: [], h: [] };
    const power = facePoints(surfaceData.faces, defaultMeshType || 0, [1, 1, 1], [1, 1, 1], -1, 0, lists);
    if (mirror) {
        const n = mirror.direction, a1 = (n + 1) % 3, a2 = (n + 2) % 3;
        const ef = [1, 1, 1], hf = [1, 1, 1];
        const t = String(mirror.type).toUpperCase();
        if (t === 'PEC') { hf[n] = -1; ef[a1] = -1; ef[a2] = -1; }
        else if (t === 'PMC') { ef[n] = -1; hf[a1] = -1; hf[a2] = -1; }
        facePoints(surfaceData.faces, defaultMeshType || 0, ef, hf, n, mirror.position, lists);
    }
    return { meta: Float32Array.from(lists.meta), accE: Float64Array.from(lists.e), accH: Float64Array.from(lists.h),
             power };
}

// computeNF2FF (src/nf2ff.mjs:102-436).  opts.engine: an initialised
// HIPEngine whose device runs it (else a context is made for the call).
export function computeNF2FF(surfaceData, freq, theta, phi, center, radius, opts) {
    center = center || [0, 0, 0];
    radius = radius === undefined ? 1 : radius;
    opts = opts || {};
    const meshType = opts.meshType || 0;
    const cxyz = meshType === 1
        ? [center[0] * Math.cos(center[1]), center[0] * Math.sin(center[1]), center[2]]
        : [center[0], center[1], center[2]];
    const pts = surfacePoints(surfaceData, meshType, opts.mirror || null);
    const nA = theta.length * phi.length;
    const out = new Float64Array(5 * nA);
    if (pts.meta.length) {
        const own = !(opts.engine && opts.engine._h);
        const h = own ? addon.create(0) : opts.engine._h;
        if (!h) throw new Error('computeNF2FF: no HIP device');
        try {
            addon.farfieldSurface(h, pts.meta, pts.accE, pts.accH, Float32Array.from(theta), Float32Array.from(phi),
                Float64Array.from(cxyz), freq, radius, out);
        } finally {
            if (own) addon.destroy(h);
        }
    }
    const E_theta_re = new Float64Array(nA), E_theta_im = new Float64Array(nA);
    const E_phi_re = new Float64Array(nA), E_phi_im = new Float64Array(nA), P_rad = new Float64Array(nA);
    let pMax = 0;
    for (let i = 0; i < nA; i++) {
        P_rad[i] = out[5 * i];
        E_theta_re[i] = out[5 * i + 1];
        E_theta_im[i] = out[5 * i + 2];
        E_phi_re[i] = out[5 * i + 3];
        E_phi_im[i] = out[5 * i + 4];
        if (P_rad[i] > pMax) pMax = P_rad[i];
    }
    const Dmax = pts.power > 0 ? pMax * 4 * Math.PI * radius * radius / pts.power : 0;
    return { E_theta_re, E_theta_im, E_phi_re, E_phi_im, Prad: pts.power, Dmax, P_rad };
}

export { C0 };
