"""Writes the golden fixtures of tests/golden from the reference's own host ops.

Run in a container where /root/reference exists and oracle/_ref was built
(`make -C oracle ref`):

    python tests/golden/make_golden.py

Fixtures are data only (inputs and the reference's outputs), stored as .npz:
  rannew64.npz        rannew64.cu:55-126 sequences (fresh process per seed)
  initial_velocity.npz set_initial_velocity.cu:7-42 on 300 clumps (fresh process)
  setup.npz           generate_particles + calculate_particle_parameters + place_particles
  steps_fill.npz      a 96-clump filling state, initial force pass + 12 steps
  steps_avalanche.npz an avalanche-phase state touching every boundary module + 4 steps
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from avalancheflume3d_amd import abi  # noqa: E402
from oracle import oracle  # noqa: E402
from tests import scenarios  # noqa: E402

SEEDS = [26752658, 12345, 987654321]


def _child(code: str) -> bytes:
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, check=True)
    return r.stdout


def rannew64() -> None:
    out = {}
    for s in SEEDS:
        code = ("import ctypes as C, numpy as np, sys\n"
                "from oracle import oracle\n"
                "lib = oracle.ref_lib()\n"
                f"d = C.c_long({s})\n"
                "v = np.array([lib.ref_rannew64(C.byref(d)) for _ in range(2000)])\n"
                "sys.stdout.buffer.write(v.tobytes())\n")
        out[f"seed_{s}"] = np.frombuffer(_child(code), dtype=np.float64)
    # re-seeding with a negative idum mid-sequence (rannew64.cu:69-81)
    code = ("import ctypes as C, numpy as np, sys\n"
            "from oracle import oracle\n"
            "lib = oracle.ref_lib()\n"
            "d = C.c_long(777)\n"
            "a = [lib.ref_rannew64(C.byref(d)) for _ in range(300)]\n"
            "d = C.c_long(-4242)\n"
            "a += [lib.ref_rannew64(C.byref(d)) for _ in range(300)]\n"
            "sys.stdout.buffer.write(np.array(a).tobytes())\n")
    out["reseed_777_then_-4242"] = np.frombuffer(_child(code), dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "rannew64.npz"), **out)


def initial_velocity() -> None:
    code = ("import numpy as np, sys\n"
            "from avalancheflume3d_amd import abi\n"
            "from oracle import oracle\n"
            "from tests import scenarios\n"
            "sys_, geo, fun, nt = scenarios.reference_system(300)\n"
            "pp = np.zeros(300, dtype=abi.particle_dtype)\n"
            "oracle.ref_lib().ref_set_initial_velocity(abi.ptr(sys_), abi.ptr(pp))\n"
            "sys.stdout.buffer.write(np.ascontiguousarray(np.concatenate([pp['vv'], pp['ww']], 1)).tobytes())\n")
    vw = np.frombuffer(_child(code), dtype=np.float64).reshape(300, 6)
    np.savez_compressed(os.path.join(HERE, "initial_velocity.npz"), vv=vw[:, :3], ww=vw[:, 3:])


def setup() -> None:
    lib = oracle.ref_lib()
    sys_, geo, fun, nt = scenarios.reference_system(1000)
    pp = np.zeros(1000, dtype=abi.particle_dtype)
    lib.ref_generate_particles(abi.ptr(pp), abi.ptr(sys_))
    lib.ref_calculate_particle_parameters(abi.ptr(sys_), abi.ptr(pp))
    lib.ref_place_particles(abi.ptr(sys_), abi.ptr(pp))
    np.savez_compressed(os.path.join(HERE, "setup.npz"), pp=pp.view(np.uint8))


def steps(name: str, inst, nsteps: int, chunks: int) -> None:
    r = oracle.RefSim(inst.sys, inst.fun, inst.ntags, inst.phase, inst.particles)
    r.initial()
    res = {"input": inst.particles.view(np.uint8), "sys": np.frombuffer(bytes(inst.sys), np.uint8),
           "fun": np.frombuffer(bytes(inst.fun), np.uint8),
           "geo": np.frombuffer(bytes(inst.geo), np.uint8),
           "meta": np.array([inst.ntags, inst.phase, nsteps, chunks])}
    g = r.get()
    res["ff_0"], res["tq_0"], res["aa_0"] = g["ff"], g["tq"], g["aa_i"]
    for k in range(1, chunks + 1):
        r.step(nsteps // chunks)
        g = r.get()
        res[f"pp_{k}"] = g["pp"].view(np.uint8)
        res[f"ff_{k}"], res[f"tq_{k}"], res[f"aa_{k}"] = g["ff"], g["tq"], g["aa_i"]
        res[f"time_{k}"] = np.array([g["sys"].time])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **res)


if __name__ == "__main__":
    assert oracle.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    rannew64()
    initial_velocity()
    setup()
    steps("steps_fill", scenarios.hopper_state(96, 3), 12, 3)
    steps("steps_avalanche", scenarios.avalanche_state(6, 9), 4, 2)
    print("fixtures written to", HERE)
