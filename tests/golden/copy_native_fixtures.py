"""Write tests/golden/dipole_openems.json and
tests/golden/engine_comparison_openems.json from the reference's native
openEMS fixture data (tests/fixtures/dipole/*.csv, tests/fixtures/
engine_comparison/probe_vp_*.csv under the reference checkout): probe
headers, sample times and the values of every column (Ex, Ey, Ez or Hx,
Hy, Hz for field probes; the voltage for the integral probe).  Data only:
the numbers openEMS v0.0.36-145-g8ee535b printed.

    python tests/golden/copy_native_fixtures.py /root/reference
"""
import json
import os
import sys


def read_csv(path):
    head, t, cols = [], [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("%"):
                head.append(line)
                continue
            v = [float(x) for x in line.split()]
            t.append(v[0])
            cols.append(v[1:])
    return {"header": head, "time_s": t, "values": cols}


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    fx = os.path.join(ref, "tests", "fixtures")
    src = "reference tests/fixtures/%s (native openEMS v0.0.36-145-g8ee535b run, generated by " \
          "tests/generate_fixtures.py:%s)"
    dip = {"source": src % ("dipole", "313-404"), "probes": {}}
    with open(os.path.join(fx, "dipole", "reference.json")) as f:
        r = json.load(f)
    dip.update({k: r[k] for k in ("drawing_unit", "f_max", "dipole_length_du", "tolerances")})
    for name in ("et1", "et2", "ht1", "ht2"):
        dip["probes"][name] = read_csv(os.path.join(fx, "dipole", "probe_%s.csv" % name))
    with open(os.path.join(here, "dipole_openems.json"), "w") as f:
        json.dump(dip, f)
    eng = {"source": src % ("engine_comparison", "439-520"), "probes": {}}
    with open(os.path.join(fx, "engine_comparison", "comparison.json")) as f:
        eng["comparison"] = json.load(f)
    for e in ("basic", "sse", "sse-compressed"):
        eng["probes"][e] = read_csv(os.path.join(fx, "engine_comparison", "probe_vp_%s.csv" % e))
    with open(os.path.join(here, "engine_comparison_openems.json"), "w") as f:
        json.dump(eng, f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
