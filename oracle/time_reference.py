"""Time the reference Fortran's CPU path in THIS container -- test
infrastructure (the CPU baseline), never shipped or run on the GPU box.

The timing build (``make -C oracle ref-timing``: the reference's own
src/2d and src/3d sources compiled by flang with -O3 -march=native and
OpenMP, linked with the oracle's Riemann solvers; output in
oracle/_ref/timing, which is not sent to the GPU box) is timed on
fixed numbers of steps of the benchmark workloads at their full sizes:

* 2D Euler 4096^2 (Roe + Harten-Hyman, MC, method 1,2,2, extrapolation):
  claw2 over N accepted steps, one thread (the reference's 2D step has no
  OpenMP: src/2d/step2.f90, flux2.f90);
* 2D acoustics pulse at 4096^2 (MC, walls): claw2, one thread;
* 3D layered acoustics 512^3 (superbee, method 22): step3 calls with
  OMP_NUM_THREADS = 1 and 8 (the reference's OpenMP loop over blocks of
  slices, src/3d/step3.f90:132-869).

Each measurement runs in a child process (its own OMP_NUM_THREADS) and
prints one JSON object; the parent prints them as a markdown table.

    python oracle/time_reference.py [--quick]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TIMING = os.path.join(HERE, "_ref", "timing")
# run as a script, this folder comes first on sys.path and would shadow the
# oracle package with oracle/oracle.py
sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, ROOT)

CASES = {
    # name: (description, threads, steps, quick steps)
    "euler2d": ("2D Euler 4096^2, claw2", (1,), 10, 1),
    "acoustics2d": ("2D acoustics pulse 4096^2, claw2", (1,), 10, 1),
    "acoustics3d": ("3D layered acoustics 512^3, step3 method 22", (1, 8), 2, 1),
}


class _Quiet:
    """fd 1 to /dev/null while the Fortran runs (it prints progress)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def _run_case(name: str, threads: int, steps: int) -> dict:
    import numpy as np

    from classic_amd import problems as P
    from oracle import oracle as O
    from oracle import oracle3 as O3

    if name in ("euler2d", "acoustics2d"):
        O._ref = C.CDLL(os.path.join(TIMING, "libclawref2d.so"))
        n = 4096
        if name == "euler2d":
            q, dx, dy = P.euler_quadrants(n, n, seed=0)
            p = O.Problem(4, 3, n, n, O.SOLVER_EULER, O.solver_params(gamma=1.4, efix=True),
                          method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4, 4], mthbc=(1, 1, 1, 1))
            dt0 = 0.9 * dx / 3.0
        else:
            q, dx, dy = P.acoustics_pulse(n, n)
            p = O.Problem(3, 2, n, n, O.SOLVER_ACOUSTICS, O.solver_params(1.0, 4.0),
                          method=(1, 2, 2, 0, 0, 0, 0), mthlim=[4, 4], mthbc=(3, 3, 3, 3))
            dt0 = 0.9 * dx / 2.0
        with _Quiet():
            t0 = time.perf_counter()
            r = O.claw2(p, q, dx, dy, 0.0, 1e9, [dt0, 1e99], [1.0, 0.9], steps, backend="ref")
            el = time.perf_counter() - t0
        done = int(r.nv[1])
        cells = n * n
    else:
        lib = C.CDLL(os.path.join(TIMING, "libclawref3d.so"))
        n = 512
        q, aux, _ = P.acoustics3d_pulse(n, seed=0)
        p = O3.Problem3(n, n, n, mthbc=(3,) * 6, mthlim=(2, 2))
        q = O3.bc3(p, q)
        aux = np.ascontiguousarray(aux)
        qnew = q.copy()
        cfl = C.c_double(0.0)
        g = p.geom
        dt = 0.9 * g.dx / 2.0
        el = 0.0
        D = C.POINTER(C.c_double)
        Ip = C.POINTER(C.c_int)
        for _ in range(steps):
            np.copyto(qnew, q)
            with _Quiet():
                t0 = time.perf_counter()
                rc = lib.ref3_step3_nt(threads, p.meqn, p.mwaves, p.maux, p.mbc, n, n, n,
                                       q.ctypes.data_as(D), qnew.ctypes.data_as(D),
                                       aux.ctypes.data_as(D), C.c_double(g.dx), C.c_double(g.dy),
                                       C.c_double(g.dz), C.c_double(dt),
                                       p.method.ctypes.data_as(Ip), p.mthlim.ctypes.data_as(Ip),
                                       0, C.byref(cfl))
                el += time.perf_counter() - t0
            if rc != 0:
                raise RuntimeError(f"step3 failed rc={rc}")
            q, qnew = O3.bc3(p, qnew), q
        done = steps
        cells = n ** 3
    return {"case": name, "threads": threads, "steps": done, "seconds": el,
            "mcell_updates_per_s": cells * done / el / 1e6}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case")
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--quick", action="store_true", help="one step per case (a check)")
    a = ap.parse_args()
    if a.case:
        print(json.dumps(_run_case(a.case, a.threads, a.steps)), flush=True)
        return
    for lib in ("libclawref2d.so", "libclawref3d.so"):
        if not os.path.exists(os.path.join(TIMING, lib)):
            raise SystemExit(f"{TIMING}/{lib} missing: run make -C oracle ref-timing")
    rows = []
    for name, (desc, threads, steps, quick) in CASES.items():
        for t in threads:
            env = dict(os.environ, OMP_NUM_THREADS=str(t))
            cmd = [sys.executable, os.path.abspath(__file__), "--case", name, "--threads", str(t),
                   "--steps", str(quick if a.quick else steps)]
            out = subprocess.run(cmd, env=env, check=True, capture_output=True, text=True).stdout
            # the Fortran runtime flushes its own messages at exit, after the JSON line
            r = json.loads([ln for ln in out.splitlines() if ln.startswith('{"case"')][-1])
            r["desc"] = desc
            rows.append(r)
            print(f"# {name} threads={t}: {r['mcell_updates_per_s']:.3f} Mcell-updates/s "
                  f"({r['steps']} steps, {r['seconds']:.1f} s)", file=sys.stderr, flush=True)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    print(f"host: {cpu}, {os.cpu_count()} logical CPUs")
    print()
    print("| workload | threads | steps timed | seconds | Mcell-updates/s |")
    print("|---|---|---|---|---|")
    for r in rows:
        print(f"| {r['desc']} | {r['threads']} | {r['steps']} | {r['seconds']:.1f} | "
              f"{r['mcell_updates_per_s']:.3f} |")


if __name__ == "__main__":
    main()
