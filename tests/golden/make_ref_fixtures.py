"""Write the fixtures that pin the cross-section and Maxwellian tables.

Runs oracle/_ref/ref_dump, which is linked from object files compiled from
the reference's own src/PhysicsData.c and src/Particle.c (see the `ref`
target of oracle/Makefile), and stores its output as data:

  ar_xsec_table.json    log10 energy grid, energies and the 7 argon cross
                        sections on it (Argon_CrossSectionSET arithmetic)
  maxwell_vsave.json    size, every 37th entry and the sum of the
                        Maxwellian inverse-CDF table (maxwellv)
  xsec_raw_checksum.json  size and sums of each tabulated data set
  o2_xsec_table.json, aro2_xsec_table.json
                        the 81 x 58 oxygen and 81 x 68 argon/oxygen tables and
                        their thresholds, every reaction switch on
                        (Oxygen_CrossSectionSET / ArO2_CrossSectionSET
                        arithmetic on the reference's own fit functions)

Only runs where the reference tree is present:
    make -C oracle ref && python tests/golden/make_ref_fixtures.py
"""
import base64
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-s"])
    with open(os.path.join(ROOT, "oracle", "_ref", "ref_dump.txt")) as f:
        lines = f.read().split("\n")
    sets, grid, ar, vs = {}, None, {}, None
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("SET "):
            _, s, n = ln.split()
            rows = [lines[i + 1 + k].split() for k in range(int(n))]
            sets[int(s)] = {"n": int(n), "sum_x": sum(float(r[0]) for r in rows),
                            "sum_y": sum(float(r[1]) for r in rows)}
            i += int(n) + 1
        elif ln.startswith("GRID "):
            n = int(ln.split()[1])
            rows = [lines[i + 1 + k].split() for k in range(n)]
            grid = ([float(r[0]) for r in rows], [float(r[1]) for r in rows])
            i += n + 1
        elif ln.startswith("AR "):
            r = int(ln.split()[1])
            ar[r] = [float(lines[i + 1 + k]) for k in range(len(grid[0]))]
            i += len(grid[0]) + 1
        elif ln.startswith("VSAVE "):
            n = int(ln.split()[1])
            vs = [float(lines[i + 1 + k]) for k in range(n)]
            i += n + 1
        else:
            i += 1
    with open(os.path.join(HERE, "ar_xsec_table.json"), "w") as f:
        json.dump({"source": "oracle/_ref/ref_dump (reference CrossSection + FUNC_decay)",
                   "xe": grid[0], "xee": grid[1], "cx": [ar[r] for r in range(7)]}, f)
    idx = list(range(0, len(vs), 37))
    import numpy as np
    with open(os.path.join(HERE, "maxwell_vsave.json"), "w") as f:
        json.dump({"source": "oracle/_ref/ref_dump (reference F)", "nvel": len(vs), "index": idx,
                   "value": [vs[k] for k in idx],
                   "sum": float(np.sum(np.array(vs, dtype=np.float32).astype(np.float64)))}, f)
    with open(os.path.join(HERE, "xsec_raw_checksum.json"), "w") as f:
        json.dump({str(k): v for k, v in sets.items()}, f, indent=0)
    # oxygen and argon/oxygen tables (float32, little-endian, base64) and thresholds
    for tag, th_tag, name in (("O2TABLE", "O2TH", "o2_xsec_table.json"), ("AO2TABLE", "AO2TH", "aro2_xsec_table.json")):
        k = next(i for i, ln in enumerate(lines) if ln.startswith(tag + " "))
        n, m = (int(v) for v in lines[k].split()[1:])
        tab = np.array([[float(v) for v in lines[k + 1 + r].split()] for r in range(n)], dtype=np.float32)
        t = lines.index(th_tag)
        th = np.array([float(v) for v in lines[t + 1].split()], dtype=np.float32)
        assert tab.shape == (n, m) and th.shape == (m,)
        with open(os.path.join(HERE, name), "w") as f:
            json.dump({"source": "oracle/_ref/ref_dump (reference O2sigma* fits, CrossSection, FUNC_decay; "
                                 "every reaction switch 1)",
                       "rows": n, "cols": m,
                       "table_f32_b64": base64.b64encode(tab.astype("<f4").tobytes()).decode(),
                       "th_e_f32_b64": base64.b64encode(th.astype("<f4").tobytes()).decode()}, f, indent=0)
    print("fixtures written")


if __name__ == "__main__":
    main()
