"""Planted mistakes in the oracle (CPU, run by hand, not collected by pytest):

    python tests/oracle_mutations.py

Each entry below alters one expression of oracle/dem_oracle.c the way a plausible slip of a
restatement would (a dropped factor, a wrong sign, another order, the wrong mass or rule),
builds the altered copy in a temporary directory and runs the CPU pins of the oracle
(tests/test_oracle_pins.py, tests/test_closed_forms.py) against it through AVF_ORACLE_LIB.
Every alteration must make at least one pin fail; the script exits non-zero otherwise.
"""
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
src = open(os.path.join(ROOT, "oracle", "dem_oracle.c")).read()

M = {
 'slide 0.8->1.0': ('aux_fric = 0.8 * mu * normal / c->dvvct_mag;', 'aux_fric = 1.0 * mu * normal / c->dvvct_mag;'),
 'wall rule = pp rule': ('int slide = pp_rule ? (fabs(ff_sdt) > mu * normal) : (fabs(ff_sdt) > mu * normal || xi == 0.0);', 'int slide = (fabs(ff_sdt) > mu * normal);'),
 'stick 2e6->1e6': ('ff_sdt = (1.0e12 * mass_f * xi + 2.0e6 * mass_f * xi_dot);', 'ff_sdt = (1.0e12 * mass_f * xi + 1.0e6 * mass_f * xi_dot);'),
 'hop km no 0.5': ('''    double km = 0.5 * kappa * mass, gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 r = orc_sphere(p, ns), rrc = {0, 0, 0}, nvec = {0, 0, 0};
      double dist, comp = 0.0;
      int touch = 0;
      if (r.z < zbottom''', '''    double km = kappa * mass, gm = 0.5 * gamma * mass;
    for (int ns = 0; ns < 4; ns++) {
      avf_double3 r = orc_sphere(p, ns), rrc = {0, 0, 0}, nvec = {0, 0, 0};
      double dist, comp = 0.0;
      int touch = 0;
      if (r.z < zbottom'''),
 'hop first wall wins': ('''          dist = xfront - r.x;
          if (dist <= ri) {''', '''          dist = xfront - r.x;
          if (dist <= ri && !touch) {'''),
 'gate open always': ('touch = (time < 0.0) ? 5 : 0;', 'touch = 5;'),
 'corner contact point sign': ('rrotc.x = q.x + ri * nvecr.x;', 'rrotc.x = q.x - ri * nvecr.x;'),
 'inter_flag ignored': ('if (!geo->inter_flag && time > 0.0) gm = mu = cns_1 = 0.0;', ';'),
 'dep normal sign': ('''      nvec.x = sind;
      nvec.y = 0.0;
      nvec.z = -cosd;''', '''      nvec.x = -sind;
      nvec.y = 0.0;
      nvec.z = -cosd;'''),
 'flume before pp': ('''    if (mods & AVF_MOD_PP) orc_pp_passes(s);
    if (mods & AVF_MOD_FLUME) orc_force_flume(s->ff, s->tq, &s->sys, s->pp);''', '''    if (mods & AVF_MOD_FLUME) orc_force_flume(s->ff, s->tq, &s->sys, s->pp);
    if (mods & AVF_MOD_PP) orc_pp_passes(s);'''),
 'hop after pp while filling': ('''    if (mods & AVF_MOD_PP) orc_pp_passes(s);
    if (mods & AVF_MOD_HOP)
      for (int nw = 0; nw < 3; nw++) orc_force_hop(s->ff, s->tq, &s->sys, s->pp, &s->geo, nw);
  } else {''', '''    if (mods & AVF_MOD_HOP)
      for (int nw = 0; nw < 3; nw++) orc_force_hop(s->ff, s->tq, &s->sys, s->pp, &s->geo, nw);
    if (mods & AVF_MOD_PP) orc_pp_passes(s);
  } else {'''),
 'cell_dim vzmin sign': ('rmin.z -= 1.05 * rgyr - vzmin * ddt;', 'rmin.z -= 1.05 * rgyr + vzmin * ddt;'),
 'cell_dim vxmax missing': ('rmax.x += 1.05 * rgyr + vxmax * ddt;', 'rmax.x += 1.05 * rgyr;'),
 'meff one clump': ('meff = 0.5 * (massm + massl);', 'meff = massl;'),
 'pp contact point radius': ('rrc.x = rr_1.x + rad_1 * nvec.x;', 'rrc.x = rr_1.x + rad_2 * nvec.x;'),
 'pairing order swapped': ('int smod = (sph_id - 1) % 4, sdiv = (sph_id - 1) / 4;', 'int sdiv = (sph_id - 1) % 4, smod = (sph_id - 1) / 4;'),
 'cells k before i': ('''    for (int idel = -1; idel <= 1; idel++)
      for (int jdel = -1; jdel <= 1; jdel++)
        for (int kdel = -1; kdel <= 1; kdel++) {''', '''    for (int kdel = -1; kdel <= 1; kdel++)
      for (int jdel = -1; jdel <= 1; jdel++)
        for (int idel = -1; idel <= 1; idel++) {'''),
}
M['frozen v kept'] = ('      p->vv.x = p->vv.y = p->vv.z = 0.0;\n      continue;', '      continue;')
M['get_accel back-rotation transposed'] = ('al.x = ab.x * Rx.x + ab.y * Rx.y + ab.z * Rx.z;', 'al.x = ab.x * Rx.x + ab.y * Ry.x + ab.z * Rz.x;')
M['gyro sign'] = ('ab.x = tb.x / I.x + Cyz * wb.y * wb.z;', 'ab.x = tb.x / I.x - Cyz * wb.y * wb.z;')
M['finish dt for hdt'] = ('vv.x += hdt * (af.x - ai.x);', 'vv.x += 2.0 * hdt * (af.x - ai.x);')
M['rotate_vectors sign'] = ('out[k].z = in[k].x * sint + in[k].z * cost;', 'out[k].z = -in[k].x * sint + in[k].z * cost;')

TESTS = ["tests/test_oracle_pins.py", "tests/test_closed_forms.py"]


def main() -> int:
    missed = 0
    with tempfile.TemporaryDirectory() as tmp:
        hdr = open(os.path.join(ROOT, "oracle", "dem_oracle.h")).read().replace(
            "../include/avflume.h", os.path.join(ROOT, "include", "avflume.h"))
        open(os.path.join(tmp, "dem_oracle.h"), "w").write(hdr)
        for k, (name, (a, b)) in enumerate(M.items()):
            if a not in src:
                print(f"{name:36s} PATTERN NOT FOUND")
                missed += 1
                continue
            c = os.path.join(tmp, f"m{k}.c")
            lib = os.path.join(tmp, f"m{k}.so")
            open(c, "w").write(src.replace(a, b, 1))
            subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                            "-I", tmp, "-o", lib, c, "-lm"], check=True)
            env = dict(os.environ, AVF_ORACLE_LIB=lib)
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p",
                                "no:cacheprovider"] + TESTS, cwd=ROOT, env=env,
                               capture_output=True, text=True)
            first = next((ln for ln in r.stdout.splitlines() if ln.startswith("FAILED")), "")
            print(f"{name:36s} {'caught: ' + first[7:100] if r.returncode else 'MISSED'}",
                  flush=True)
            missed += r.returncode == 0
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main())
