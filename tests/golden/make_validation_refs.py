"""Writes tests/golden/validation_refs.json from the reference's validation data.

Sources (data files of tumaer/jax-sph, read as text):
* validation/ref/tgv3d_ref_100.txt: 3D Taylor-Green vortex at Re = 100 from
  JAX-Fluids (Nx = 64), columns t, -dE/dt, E (E = kinetic energy / volume);
  every 10th row is kept.
* validation/ref/ldc_data_u_vel.csv, ldc_data_v_vel.csv: Ghia et al. (1982)
  lid-driven cavity centreline velocities; the Re = 100 columns are kept
  (the u file's "0,00000" entry is read as 0, as validate.py:404-406 does).
* validation/ref/rpf_t_{2,10,20,100}.csv: reverse Poiseuille flow (cases/rpf.py)
  velocity profiles u(y) over one half-channel, y in [0, 1]; the number in the
  file name is ten times the time (the profiles are those of t = 0.2, 1, 2
  and 10 for g = 1 and nu = 0.1, as the closed-form series shows).
* validation/ref/dambreak_data.py: Martin & Moyce (1952) surge-front data
  (arrays time_front, y_front: front distance from the wall over the column
  side, Z, against T = t sqrt(2 g / a)), read as text.

Run from the repository root: python tests/golden/make_validation_refs.py
"""
import csv
import json
import os
import re

REF = "/root/reference/validation/ref"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "validation_refs.json")


def _ghia(path, axis):
    rows = list(csv.reader(open(path)))
    head = rows[1]
    i_ax, i_100 = head.index(axis), head.index("100")
    pos, val = [], []
    for r in rows[2:]:
        if len(r) <= i_100 or not r[i_ax].strip():
            continue
        pos.append(float(r[i_ax]))
        val.append(float(r[i_100].replace(",", ".")))
    return pos, val


def main():
    t, dedt, e = [], [], []
    for k, line in enumerate(open(os.path.join(REF, "tgv3d_ref_100.txt"))):
        if k % 10 == 9:
            a, b, c = (float(x) for x in line.split(","))
            t.append(a), dedt.append(b), e.append(c)
    yu, u = _ghia(os.path.join(REF, "ldc_data_u_vel.csv"), "y")
    xv, v = _ghia(os.path.join(REF, "ldc_data_v_vel.csv"), "x")
    rpf = {"source": "validation/ref/rpf_t_{2,10,20,100}.csv (t = number / 10)"}
    for num in (2, 10, 20, 100):
        rows = list(csv.reader(open(os.path.join(REF, f"rpf_t_{num}.csv"))))[1:]
        rpf[str(num / 10)] = {"y": [float(r[0]) for r in rows if r],
                              "u": [float(r[1]) for r in rows if r]}
    src = open(os.path.join(REF, "dambreak_data.py")).read()

    def arr(name):
        m = re.search(name + r" = np.array\(\s*\[([^\]]*)\]", src)
        return [float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]
    mm = {"source": "validation/ref/dambreak_data.py (time_front, y_front)",
          "T": arr("time_front"), "Z": arr("y_front")}
    json.dump({"db_front_martin_moyce": mm, "rpf_profiles": rpf,
               "tgv3d_re100": {"source": "validation/ref/tgv3d_ref_100.txt (every 10th row)",
                               "t": t, "minus_dEdt": dedt, "E": e},
               "ldc_re100_ghia": {"source": "validation/ref/ldc_data_{u,v}_vel.csv, Re=100",
                                  "y": yu, "u": u, "x": xv, "v": v}},
              open(OUT, "w"), indent=1)
    print("wrote", OUT, len(t), "TGV rows,", len(u), "u and", len(v), "v points")


if __name__ == "__main__":
    main()
