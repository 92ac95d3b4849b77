"""Writes tests/golden/validation_refs.json from the reference's validation data.

Sources (data files of tumaer/jax-sph, read as text):
* validation/ref/tgv3d_ref_100.txt: 3D Taylor-Green vortex at Re = 100 from
  JAX-Fluids (Nx = 64), columns t, -dE/dt, E (E = kinetic energy / volume);
  every 10th row is kept.
* validation/ref/ldc_data_u_vel.csv, ldc_data_v_vel.csv: Ghia et al. (1982)
  lid-driven cavity centreline velocities; the Re = 100 columns are kept
  (the u file's "0,00000" entry is read as 0, as validate.py:404-406 does).

Run from the repository root: python tests/golden/make_validation_refs.py
"""
import csv
import json
import os

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
    json.dump({"tgv3d_re100": {"source": "validation/ref/tgv3d_ref_100.txt (every 10th row)",
                               "t": t, "minus_dEdt": dedt, "E": e},
               "ldc_re100_ghia": {"source": "validation/ref/ldc_data_{u,v}_vel.csv, Re=100",
                                  "y": yu, "u": u, "x": xv, "v": v}},
              open(OUT, "w"), indent=1)
    print("wrote", OUT, len(t), "TGV rows,", len(u), "u and", len(v), "v points")


if __name__ == "__main__":
    main()
