"""Served-stream workload (config 5) and its driver.

A stream of independent plasticity problems -- plate-with-hole or graded
ring, uniformly at random; mesh size log-uniform in quads; hole radius
U(0.05, 0.3); yield U(200, 400) MPa; linear (H = E/100) or Swift power-law
hardening -- arrives by a Poisson process and is served by the library's
continuous-batching loop (cp_server_*).  Plates are loaded to 1.25 sigma_Y on
the far edge, rings (radially graded to an overall element ratio 1.05^59) to
1.2 sigma_Y ln(b/a) internal pressure.
"""
import math
import time

import numpy as np

from . import abi as A

E_MPA = 2.1e5


def make_problem(rng, qmin, qmax):
    quads = math.exp(rng.uniform(math.log(qmin), math.log(qmax)))
    kind = "plate" if rng.uniform() < 0.5 else "ring"
    sY = rng.uniform(200.0, 400.0)
    law = A.LAW_LINEAR if rng.uniform() < 0.5 else A.LAW_SWIFT
    mat = A.material(E_MPA, 0.3, sY, law, K_T=E_MPA / 101, n=0.2)
    if kind == "plate":
        Nr = max(4, int(round(math.sqrt(quads / 2.0))))
        Nt = 2 * Nr
        r = rng.uniform(0.05, 0.3)
        mesh = A.gen_quarter_plate(1.0, r, Nt, Nr)
        loads = A.Loads(mesh).force(1, 1.25 * sY, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
        desc = {"kind": kind, "Nt": Nt, "Nr": Nr, "r": r}
    else:
        Nr = max(4, int(round(math.sqrt(quads / 4.0))))
        Nt = 4 * Nr
        # radial grading: the 16700x60 ring's overall ratio 1.05^59 whatever
        # the radial count (1.05 per element over 350 elements would be 3e7)
        ratio = 1.05 ** (59.0 / max(1, Nr - 1))
        mesh = A.gen_quarter_ring(0.5, 1.0, Nt, Nr, ratio)
        p = 1.2 * sY * math.log(2.0)
        loads = A.Loads(mesh).force(0, -p, 0.0, norm_orient=True).fix_vert_axis(0.0).fix_hor_axis(0.0)
        desc = {"kind": kind, "Nt": Nt, "Nr": Nr, "ratio": ratio}
    desc.update({"quads": mesh.ne, "sigmaY": sY, "law": "linear" if law == A.LAW_LINEAR else "swift"})
    return mesh, loads, mat, desc


def lattice(desc):
    """Node lattice of a generated plate or ring (multigrid CG)."""
    return desc["Nt"] + 1, desc["Nr"] + 1


def make_stream(n, seed=20261203, qmin=2e3, qmax=5e5):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [make_problem(rng, qmin, qmax) for _ in range(n)]


def serve(ctx, stream, nsteps, opt, rate, hbm_budget, max_slots=16, batch_cg=True, seed=7, log=None, batch_mg=False,
          stats=None, batch_mg_max_nodes=0):
    """Submit `stream` with Poisson(rate) arrivals (rate in problems/s on the
    server clock; 0: all at once), run the loop until drained.  Returns
    per-request records; `stats` (a dict), if given, receives the server's
    lockstep batch counts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate, size=len(stream)) if rate > 0 else np.zeros(len(stream))
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = A.Server(ctx, hbm_budget, max_slots=max_slots, batch_cg=batch_cg, batch_mg=batch_mg,
                   batch_mg_max_nodes=batch_mg_max_nodes)
    ids = [None] * len(stream)
    t_start = srv.clock()
    i = 0
    last_log = time.time()
    while True:
        now = srv.clock() - t_start
        while i < len(stream) and arrivals[i] <= now:
            mesh, loads, mat, desc = stream[i]
            ids[i] = srv.submit(mesh, loads, mat, nsteps, 1.0, opt,
                                grid=lattice(desc) if opt.precond == 2 else None)
            i += 1
        in_flight = srv.step()
        if in_flight == 0:
            if i == len(stream):
                break
            time.sleep(max(0.0, min(0.01, arrivals[i] - (srv.clock() - t_start))))
        if log and time.time() - last_log > 20:
            log(f"  served {i} submitted, {in_flight} in flight, t={srv.clock() - t_start:.1f}s")
            last_log = time.time()
    if stats is not None:
        stats["batches"], stats["members"], stats["lockstep_iterations"] = srv.batch_stats()
    recs = []
    for k, rid in enumerate(ids):
        code, info, u, sig = srv.result(rid)
        if code != A.CP_OK and log:
            log(f"  request {k} failed: {srv.last_message}")
        recs.append({"code": code, "message": srv.last_message, "latency": info.t_done - info.t_submit, "queue": info.t_admit - info.t_submit,
                     "t_done": info.t_done - t_start, "cg_iters": info.cg_iters, "outer": info.outer_iters,
                     "u": u, "sig": sig})
    srv.close()
    return recs
