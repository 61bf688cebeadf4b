"""openEMS XML documents for the front-end tests, written in the format the
reference's writers use (app/ems-api.mjs GenerateXML, 847-975, and
tests/generate_fixtures.py): the patch-antenna example as the script API
would write it (app/examples.mjs:17-103 through AddLumpedPort,
app/ems-api.mjs:607-666, and CreateNF2FFBox), and the cavity and coax
fixture simulations (tests/generate_fixtures.py:65-130, 181-310), whose
native openEMS results the reference holds.  Test data only."""
import numpy as np


def _num(v):
    return repr(float(v))


def _csv(a):
    return ",".join(_num(v) for v in a)


def _pt(tag, p):
    return '<%s X="%s" Y="%s" Z="%s"/>' % (tag, _num(p[0]), _num(p[1]), _num(p[2]))


def _box(p1, p2, prio=0):
    return '<Box Priority="%d">%s%s</Box>' % (prio, _pt("P1", p1), _pt("P2", p2))


def _doc(nts, f0, fc, bc, unit, lines, props, end="1e-4", bc_extra=""):
    return ('<?xml version="1.0" encoding="UTF-8"?>\n<openEMS>\n'
            '  <FDTD NumberOfTimesteps="%d" endCriteria="%s" f_max="%s">\n'
            '    <Excitation Type="0" f0="%s" fc="%s"/>\n'
            '    <BoundaryCond %s%s/>\n  </FDTD>\n'
            '  <ContinuousStructure CoordSystem="0">\n    <RectilinearGrid DeltaUnit="%s">\n'
            '      <XLines>%s</XLines>\n      <YLines>%s</YLines>\n      <ZLines>%s</ZLines>\n'
            '    </RectilinearGrid>\n    <Properties>\n%s    </Properties>\n  </ContinuousStructure>\n</openEMS>\n'
            % (nts, end, _num(f0 + fc), _num(f0), _num(fc),
               " ".join('%s="%s"' % (k, v) for k, v in zip(("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"), bc)),
               bc_extra, _num(unit), _csv(lines[0]), _csv(lines[1]), _csv(lines[2]), props))


def _prop(tag, attrs, body):
    a = "".join(' %s="%s"' % (k, v) for k, v in attrs)
    return "      <%s%s>\n%s      </%s>\n" % (tag, a, body, tag)


def _prims(*prims):
    return "        <Primitives>\n" + "".join("          %s\n" % p for p in prims) + "        </Primitives>\n"


def patch_xml(w, feed_r=50.0, nf2ff_freq=2.4e9):
    """The patch example of efs_amd.workloads.patch_antenna(w's defaults) as
    the reference's script API writes it."""
    xs, ys, zs = w["lines_mm"]
    W, L = w["patch"]
    hw, hl, hpw, hpl, h = 30.0, 30.0, W / 2, L / 2, 1.524
    eps_r = 3.38
    kappa = 1e-3 * 2 * np.pi * 2.45e9 * 8.854187817e-12 * eps_r
    feed = [-6.0, 0.0, 0.0]
    stop = [-6.0, 0.0, h]
    # current probe box: the port's middle along z, two lines out on x and y
    it_lo, it_hi = list(feed), list(stop)
    for i, lines in ((0, xs), (1, ys)):
        k = int(np.argmin(np.abs(lines - feed[i])))
        it_lo[i], it_hi[i] = lines[k - 2], lines[k + 2]
    it_lo[2] = it_hi[2] = h / 2
    ins = 3
    lo = [xs[ins], ys[ins], zs[ins]]
    hi = [xs[-1 - ins], ys[-1 - ins], zs[-1 - ins]]
    faces = [([lo[0], lo[1], lo[2]], [lo[0], hi[1], hi[2]]), ([hi[0], lo[1], lo[2]], [hi[0], hi[1], hi[2]]),
             ([lo[0], lo[1], lo[2]], [hi[0], lo[1], hi[2]]), ([lo[0], hi[1], lo[2]], [hi[0], hi[1], hi[2]]),
             ([lo[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]]), ([lo[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]])]
    props = ""
    props += _prop("Metal", [("Name", "ground")], _prims(_box([-hw, -hl, 0], [hw, hl, 0], 10)))
    props += _prop("Material", [("Name", "substrate")],
                   '        <Property Epsilon="%s" Kappa="%s"/>\n' % (_num(eps_r), _num(kappa))
                   + _prims(_box([-hw, -hl, 0], [hw, hl, h], 5)))
    props += _prop("Metal", [("Name", "patch")], _prims(_box([-hpw, -hpl, h], [hpw, hpl, h], 10)))
    props += _prop("LumpedElement", [("Name", "port_resist_1"), ("Direction", "2"), ("R", _num(feed_r)),
                                     ("C", "0"), ("L", "0")], _prims(_box(feed, stop, 5)))
    props += _prop("Excitation", [("Name", "port_excite_1"), ("Type", "0"), ("Excite", "0,0,1")],
                   _prims(_box(feed, stop, 5)))
    props += _prop("ProbeBox", [("Name", "port_ut1"), ("Type", "0"), ("Weight", "-1")], _prims(_box(feed, stop, 5)))
    props += _prop("ProbeBox", [("Name", "port_it1"), ("Type", "1"), ("Weight", "1"), ("NormDir", "2")],
                   _prims(_box(it_lo, it_hi, 5)))
    for fld, dt in (("E", 10), ("H", 11)):
        props += _prop("DumpBox", [("Name", "nf2ff_" + fld), ("DumpType", str(dt)), ("DumpMode", "1"),
                                   ("FileType", "1")],
                       "        <FD_Samples>%s</FD_Samples>\n" % _num(nf2ff_freq)
                       + _prims(*[_box(a, b, 0) for a, b in faces]))
    return _doc(w["nTS"], w["f0"], w["fc"], ["2"] * 6, 1e-3, [xs, ys, zs], props)


def cavity_xml(cav):
    """The cavity fixture's simulation (50 x 20 x 60 mm PEC box, soft source
    on a curve from the node at 2/3 of each axis to the next node, V_z
    probe)."""
    d = cav["dimensions_m"]
    n = cav["mesh_points"]
    mx, my, mz = (np.linspace(0, d[k], n[a]) for k, a in (("a", "x"), ("b", "y"), ("d", "z")))
    ex = [len(mx) * 2 // 3, len(my) * 2 // 3, len(mz) * 2 // 3]
    px, py, pz = mx[len(mx) // 4], my[len(my) // 2], len(mz) // 5
    props = _prop("Excitation", [("Name", "excite1"), ("Type", "0"), ("Excite", "1,1,1")],
                  _prims('<Curve Priority="0">%s%s</Curve>' % (
                      _pt("Vertex", [mx[ex[0]], my[ex[1]], mz[ex[2]]]),
                      _pt("Vertex", [mx[ex[0] + 1], my[ex[1] + 1], mz[ex[2] + 1]]))))
    props += _prop("ProbeBox", [("Name", "ut1z"), ("Type", "0"), ("Weight", "1"), ("NormDir", "-1")],
                   _prims(_box([px, py, mz[pz]], [px, py, mz[pz + 1]])))
    return _doc(20000, 5.5e9, 4.5e9, ["0"] * 6, 1.0, [mx, my, mz], props, end="1e-6")


def coax_xml(res=20.0, length=200.0, ri=100.0, rai=230.0, raa=240.0):
    """The coax fixture's simulation: a 200 mm air line, PEC inner cylinder
    and outer shell, TEM soft source with 1/rho weights at z = 0, PML_8 at
    z max, voltage and current probes at mid-length."""
    x_min = -2.5 * res - raa
    mx = [x_min + res * k for k in range(int(round((2 * (raa + 2.5 * res)) / res)) + 1)]
    my = list(mx)
    mz = np.linspace(0, length, int(length / res) + 1)
    mid_z = length / 2
    cur = ri + 3 * res
    probe_y = my[len(my) // 2]
    props = _prop("Metal", [("Name", "PEC")], _prims(
        '<Cylinder Priority="1" Radius="%s">%s%s</Cylinder>' % (_num(ri), _pt("P1", [0, 0, 0]),
                                                                 _pt("P2", [0, 0, length])),
        '<CylindricalShell Priority="0" Radius="%s" ShellWidth="%s">%s%s</CylindricalShell>' % (
            _num(0.5 * (raa + rai)), _num(raa - rai), _pt("P1", [0, 0, 0]), _pt("P2", [0, 0, length]))))
    props += _prop("Excitation", [("Name", "excite"), ("Type", "0"), ("Excite", "1,1,0")],
                   '        <Weight X="x/(x*x+y*y)" Y="y/(x*x+y*y)" Z="0"/>\n' + _prims(
                       '<CylindricalShell Priority="0" Radius="%s" ShellWidth="%s">%s%s</CylindricalShell>' % (
                           _num(0.5 * (ri + rai)), _num(rai - ri), _pt("P1", [0, 0, 0]), _pt("P2", [0, 0, res / 2]))))
    props += _prop("ProbeBox", [("Name", "ut1"), ("Type", "0"), ("Weight", "1"), ("NormDir", "-1")],
                   _prims(_box([ri, probe_y, mid_z], [rai, probe_y, mid_z])))
    props += _prop("ProbeBox", [("Name", "it1"), ("Type", "1"), ("Weight", "1"), ("NormDir", "-1")],
                   _prims(_box([-cur, -cur, mid_z], [cur, cur, mid_z])))
    return _doc(2000, 0.0, 1e9, ["0", "0", "0", "0", "0", "PML_8"], 1e-3, [mx, my, mz], props, end="1e-6")


def dipole_xml(du=1e-6, f_max=1e9):
    """The dipole fixture's simulation: an infinitesimal dipole (hard E
    source on a two-cell z curve, lambda/50 long at f_max) at the centre of
    a cube of +-20 dipole lengths meshed at half a dipole length, Mur on
    all six faces, 500 steps; E and H point probes at +-4.5 half-lengths on
    the x axis, snapped to the mesh as the fixture's generator snaps them
    (nearest line, the first of two equally near)."""
    c0 = 299792458.0
    L = c0 / f_max / du / 50
    h = L / 2
    lines = []
    v = -20 * L
    while v <= 20 * L + 0.001:
        lines.append(v)
        v += h

    def near(val):
        return min(lines, key=lambda m: abs(m - val))
    s = 4.5 * h
    props = _prop("Excitation", [("Name", "infDipole"), ("Type", "1"), ("Excite", "0,0,1")],
                  _prims('<Curve Priority="1">%s%s</Curve>' % (_pt("Vertex", [0, 0, -h]), _pt("Vertex", [0, 0, h]))))
    for name, x, t in (("et1", -s, 2), ("et2", s, 2), ("ht1", -s, 3), ("ht2", s, 3)):
        p = [near(x), near(0.0), near(0.0)]
        props += _prop("ProbeBox", [("Name", name), ("Type", str(t)), ("Weight", "1"), ("NormDir", "-1")],
                       _prims(_box(p, p)))
    return _doc(500, 0.0, f_max, ["2"] * 6, du, [lines, lines, lines], props, end="1e-20")


def engine_comparison_xml():
    """The engine-comparison fixture's simulation: the cavity of cavity_xml
    on a coarser 16 x 8 x 18 mesh, soft source on the one-cell diagonal
    from node (10, 5, 12), V_z probe from node (5, 3, 4) to (5, 3, 5), 500
    steps."""
    mx, my, mz = np.linspace(0, 5e-2, 16), np.linspace(0, 2e-2, 8), np.linspace(0, 6e-2, 18)
    props = _prop("Excitation", [("Name", "exc"), ("Type", "0"), ("Excite", "1,1,1")],
                  _prims('<Curve Priority="0">%s%s</Curve>' % (_pt("Vertex", [mx[10], my[5], mz[12]]),
                                                               _pt("Vertex", [mx[11], my[6], mz[13]]))))
    props += _prop("ProbeBox", [("Name", "vp"), ("Type", "0"), ("Weight", "1"), ("NormDir", "-1")],
                   _prims(_box([mx[5], my[3], mz[4]], [mx[5], my[3], mz[5]])))
    return _doc(500, 5.5e9, 4.5e9, ["0"] * 6, 1.0, [mx, my, mz], props, end="1e-20")
