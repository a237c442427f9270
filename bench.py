"""Benchmark: FSSH trajectory-steps/s on one MI355X, batched (persistent, FP32 nuclei /
FP64 electronics) vs the plain GPU port (one launch per step, FP64), with the final-
population error of the batched form against the FP64 path.

Workload (one "step" of this benchmark = one full ensemble pass): 1e6 trajectories
mixing the three Tully models (single avoided crossing, dual crossing, extended
coupling), p0 on the 5..35 a.u. grid, x0 ~ N(-5, p0/(20 sqrt 2)), m = 2000, dt = 1 a.u.,
t_max = 10000 a.u., terminated at |x| > 10 (uneven lengths).  Initial conditions are
sampled on the device before the timed region; the timed region is the trajectory
integration.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nqcdynamics.jl_amd"))

METRIC = "FSSH trajectory-steps/s on 1 GPU, learned/batched vs plain at ≤0.01 pop. error"
KEY = 0x4E5143444D493335
TULLY = [(0, [0.01, 1.6, 0.005, 1.0]), (1, [0.1, 0.28, 0.015, 0.06, 0.05]), (2, [6e-4, 0.1, 0.9])]
MASS = 2000.0
NTRAJ = 1_000_000
TMAX = 10000.0
XTERM = 10.0
GRID = [-5.0, 5.0, 31]
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak
HBM_PEAK_GBS = 8000.0
# FP64 flops (an FMA counts 2) of one trajectory-step of the trace-fixed two-state formulation,
# counted from nqcd_device.h / fssh2.hip against the compiled loop: 4 Tsit5 sub-steps x 180 (63
# stage-combination FMA, 6 right-hand sides, coefficient set-up) + 185 for the last, clipped
# sub-step + 8 for the products with h formed once per step = 913, plus ~75 for the eigen, the
# coupling rotation, the interpolation set-up and the hop test.  The FP32 model and Verlet work are
# not counted.
FLOP_PER_STEP = 988.0
# Threads of the CPU baseline: the host share of one GPU on the measurement machine (16 cores)
CPU_THREADS = 16


def _ensemble(L, precision, schedule):
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_FSSH, 2, 1, 1, [MASS], 1.0, precision=precision, schedule=schedule)
    for slot, (kind, p) in enumerate(TULLY):
        E.set_model(slot, kind, p)
    E.sample_initial(L.SAMPLER_TULLY_GRID, NTRAJ, GRID, KEY, n_models=3)
    return E


def _final_pops(L, E):
    fin = E.reduce(L.OUT_FINAL_ADIABATIC_POPULATION, L.RED_MEAN, 2)
    scat = E.reduce(L.OUT_SCATTERING_1D, L.RED_MEAN, 4)
    return np.concatenate([fin, scat])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(L, gpu_states, nthreads, samples):
    """FP64 CPU restatement of the reference algorithm (the oracle) on a stratified sample of the headline
    trajectories, on `nthreads` host threads and on one core, and the error of every other reported GPU
    result against the oracle on a sample of that workload (`samples`: per-workload GPU outputs and inputs)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    stride = 13  # every 13th trajectory id: covers every model and momentum of the grid
    ids = np.arange(0, NTRAJ, stride)
    p0 = GRID[1] + (ids // 3) % GRID[2]
    r0, v0 = O.sample_gaussian(len(ids), 1, GRID[0], (p0 / (20 * np.sqrt(2)))[:, None], (p0 / MASS)[:, None], 0.0,
                               KEY, traj_ids=ids)
    sig0 = np.zeros((len(ids), 2, 2), complex); sig0[:, 0, 0] = 1
    models = [O.make_model(k, 2, 1, p) for k, p in TULLY]

    def run(sel, threads):
        t0 = time.perf_counter()
        out = O.run_ensemble(0, models, MASS, r0[sel], v0[sel], sig0[sel], state0=np.zeros(len(sel), np.int32),
                             model_id=(ids[sel] % 3).astype(np.int32), dt=1.0, nsteps=int(TMAX), key=KEY, term_x=XTERM,
                             nthreads=threads, traj_ids=ids[sel])
        return out, time.perf_counter() - t0

    out, el = run(np.arange(len(ids)), nthreads)
    states = out["state"]; steps = int(out["steps"].sum())
    one = np.arange(0, len(ids), 16)  # a sixteenth of the sample on one core
    out1, el1 = run(one, 1)
    steps1 = int(out1["steps"].sum())
    agree = float(np.mean(states == gpu_states[ids]))
    pc = np.bincount(states, minlength=2) / len(ids)
    pg = np.bincount(gpu_states[ids], minlength=2) / len(ids)
    res = {"value": steps / el, "unit": "trajectory-steps/s", "cores": nthreads, "kind": "port",
           "sample": f"{len(ids)} trajectories (every {stride}th id of the 1e6 ensemble, all lengths, "
                     f"{steps} steps, {el:.1f} s) of the FP64 CPU restatement of the reference algorithm",
           "cpu_model": _cpu_model(), "host_cpus_visible": os.cpu_count(),
           "one_core": {"value": steps1 / el1, "cores": 1,
                        "sample": f"{len(one)} of those trajectories, {steps1} steps, {el1:.1f} s"},
           "pop_err_fp32_gpu_vs_fp64_oracle_sample": float(np.abs(pc - pg).max()),
           "pop_sigma_sample": float(np.sqrt(np.max(pc * (1 - pc)) / len(ids))),
           "state_agreement_sample": agree}
    checks = {}
    for name, fn in (("spin_boson", _check_spin_boson), ("ring_polymer", _check_ring_polymer), ("iesh", _check_iesh),
                     ("learned", _check_learned), ("mapping", _check_mapping), ("nstate", _check_nstate),
                     ("cme", _check_cme)):
        if name in samples:
            t0 = time.perf_counter()
            checks[name] = fn(O, samples[name], nthreads)
            checks[name]["oracle_s"] = time.perf_counter() - t0
            _progress(f"oracle check {name}: {checks[name]}")
    res["errors_vs_fp64_oracle"] = checks
    return res


def _check_spin_boson(O, S, nthreads):
    """Diabatic population difference <sigma_z(t)> of the GPU run (FP32 bath) on its first trajectories
    against the oracle on the same draws; max over the save slots, with the sample's statistical sigma."""
    out = {}
    m = O.make_model(O.MODEL_SPIN_BOSON, 2, len(S["w"]), [1.0, 1.0], omega=S["w"], cbath=S["c"])
    n = S["n"]
    r0, v0 = O.sample_gaussian(n, len(S["w"]), 0.0, S["sdr"][None, :], 0.0, S["sdv"][None, :], KEY)
    sig0 = O.diabatic_to_adiabatic(m, r0, 0)
    for name, method in (("ehrenfest", 1), ("fssh", 0)):
        c = O.run_ensemble(method, m, 1.0, r0, v0, sig0, state0=None if method == 1 else np.full(n, -1, np.int32),
                           dt=S["dt"], nsteps=S["nsteps"], save_every=S["save_every"], nsave=S["nsave"], key=KEY,
                           nthreads=nthreads)
        szc = c["pop_di"][:, :, 0] - c["pop_di"][:, :, 1]
        szg = S[name][:, :, 0] - S[name][:, :, 1]
        out[name] = {"trajectories": n, "max_abs_err_sigma_z": float(np.abs(szg.mean(0) - szc.mean(0)).max()),
                     "sigma": float(szc.std(0).max() / np.sqrt(n))}
    return out


def _check_ring_polymer(O, S, nthreads):
    """RPSH at P = 8 (FP32 beads, FP64 electronics) on its first ring polymers against the FP64 oracle from
    the same PIMC initial beads: final adiabatic population and centroid error."""
    m = O.make_model(O.MODEL_DW_GAUSSIAN, 2, 1, S["prm"])
    n = len(S["r0"])
    sig0 = np.zeros((n, 2, 2), complex); sig0[:, 0, 0] = 1
    c = O.rp_run_ensemble(O.RP_FSSH, m, S["mass"], S["r0"], S["v0"], sig0, np.zeros(n, np.int32), dt=1.0,
                          nsteps=S["nsteps"], temperature=S["T"], save_every=S["save_every"], nsave=S["nsave"], key=KEY,
                          nthreads=nthreads)
    pc = np.mean(c["state"] == 1)
    pg = np.mean(S["state"] == 1)
    return {"workload": "rpsh_P8", "ring_polymers": n, "excited_population": {"gpu": float(pg), "oracle": float(pc)},
            "pop_error": float(abs(pg - pc)), "pop_sigma": float(np.sqrt(max(pc * (1 - pc), 1 / n) / n)),
            "state_agreement": float(np.mean(c["state"] == S["state"])),
            "max_centroid_err": float(np.abs(c["centroid"] - S["centroid"]).max())}


def _check_iesh(O, S, nthreads):
    """IESH at N = 41 (FP64 throughout) on its first trajectories against the oracle: impurity population
    per save slot."""
    md = O.na_model(M=S["M"])
    n = S["n"]
    r0, v0 = O.sample_gaussian(n, 1, 8.0, 1.0, 0.0, 1e-2, KEY)
    occ0 = np.tile(np.arange(S["ne"], dtype=np.int32), (n, 1))
    c = O.iesh_run_ensemble(0, md, S["ne"], 2000.0, r0[:, 0], v0[:, 0], occ0, dt=2.0, nsteps=S["nsteps"], key=KEY,
                            save_every=S["save_every"], nthreads=nthreads)
    return {"workload": f"iesh_N{S['M'] + 1}", "trajectories": n,
            "max_abs_err_impurity_population": float(np.abs(c["pop_di"][:, :, 0] - S["pop_imp"]).max()),
            "hop_sequences_equal": bool(np.array_equal(c["nhops"], S["nhops"]))}


def _check_learned(O, S, nthreads):
    """FSSH on the trained surrogate (and on the analytic model, FP32 nuclei) for the first trajectories of
    the 1e6-trajectory diatom workload against the FP64 oracle on the analytic model."""
    n = len(S["r0"])
    m = O.make_model(O.MODEL_DIATOM_SURFACE, 2, 6, S["params"])
    sig0 = np.zeros((n, 2, 2), complex); sig0[:, 0, 0] = 1
    c = O.run_ensemble(0, m, np.repeat(S["mass"], 3), S["r0"], S["v0"], sig0, state0=np.zeros(n, np.int32), dt=2.0,
                       nsteps=S["nsteps"], key=KEY, term_x=S["term"], term_dof=2, nthreads=nthreads)
    pc = float(np.mean(c["state"] == 1))
    ps, pa = float(np.mean(S["surrogate_state"] == 1)), float(np.mean(S["analytic_state"] == 1))
    return {"trajectories": n, "excited_population": {"oracle": pc, "surrogate_gpu": ps, "analytic_gpu": pa},
            "pop_error_surrogate": abs(ps - pc), "pop_error_analytic_fp32": abs(pa - pc),
            "pop_sigma": float(np.sqrt(max(pc * (1 - pc), 1 / n) / n))}


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def iesh_extra(L):
    """IESH on the Newns-Anderson model (MiaoSubotnik, W = 3 Gamma, half filling) at N = 41 and 101
    states: trajectory-steps/s of the one-trajectory-per-workgroup kernel and its FP64 rate."""
    res = {}
    W = 6 * 6.4e-3 / 2
    sample = None
    for M, ntraj, nsteps in ((40, 16384, 200), (100, 4096, 100)):
        E = L.Ensemble(0)
        E.set_problem(L.METHOD_IESH, M + 1, 1, 1, [2000.0], 2.0, algorithm=L.ALG_VERLET_ELECTRONICS, n_electrons=M // 2)
        E.set_model(0, L.MODEL_NEWNS_ANDERSON, [2000.0, 2e-4, 20.6097, -3.8e-3, 6.4e-3, -W, W, M])
        E.sample_initial(L.SAMPLER_GAUSSIAN, ntraj, [8.0, 1.0, 0.0, 1e-2], KEY)
        E.run(0.0, 2.0 * 10, KEY)
        st = E.run(0.0, 2.0 * nsteps, KEY, save_every=10)
        if M == 40:  # the first 64 trajectories for the oracle check in the CPU-baseline leg
            n = 64
            sample = {"M": M, "ne": M // 2, "n": n, "nsteps": nsteps, "save_every": 10,
                      "pop_imp": E.download(L.FIELD_SAVE_POP_DI)[:n, :, 0].copy(), "nhops": E.download(L.FIELD_NHOPS)[:n]}
        E.close()
        N, Ne = M + 1, M // 2
        flop = 8 * 4 * N * N * Ne + 16 * N * M + (8 / 3) * Ne ** 3  # Taylor (K ~ 8) + eigen + Gauss-Jordan
        rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
        res[f"N{N}"] = {"ntraj": ntraj, "nsteps": nsteps, "kernel_ms": st["kernel_ms"], "traj_steps_per_s": rate,
                        "flop_per_traj_step": flop, "achieved_TFLOPs": flop * rate / 1e12,
                        "frac_fp64_peak": flop * rate / 1e12 / FP64_PEAK_TFLOPS, "grid": st["grid"], "block": st["block"]}
        _progress(f"iesh N={N}: {rate:.3e} trajectory-steps/s")
    # ring-polymer IESH (BCBWavefunction), 8 beads at 300 K: trajectory-steps/s (a ring polymer = one trajectory)
    M, P, ntraj, nsteps = 40, 8, 4096, 100
    rng = np.random.default_rng(1)
    rb = rng.normal(8.0, 1.0, (ntraj, 1)) + rng.normal(0.0, 0.05, (ntraj, P))
    vb = rng.normal(0.0, 1e-2, (ntraj, 1)) + rng.normal(0.0, 1e-3, (ntraj, P))
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_IESH, M + 1, 1, 1, [2000.0], 2.0, algorithm=L.ALG_BCB_WAVEFUNCTION, nbeads=P,
                  temperature=3.166811563e-6 * 300.0, n_electrons=M // 2)
    E.set_model(0, L.MODEL_NEWNS_ANDERSON, [2000.0, 2e-4, 20.6097, -3.8e-3, 6.4e-3, -W, W, M])
    E.load_initial_iesh(rb, vb, np.tile(np.arange(M // 2, dtype=np.int32), (ntraj, 1)))
    E.run(0.0, 2.0 * 10, KEY)
    st = E.run(0.0, 2.0 * nsteps, KEY, save_every=10)
    E.close()
    rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
    res[f"N{M + 1}_P{P}"] = {"ntraj": ntraj, "nsteps": nsteps, "nbeads": P, "kernel_ms": st["kernel_ms"],
                              "traj_steps_per_s": rate, "grid": st["grid"], "block": st["block"]}
    _progress(f"ring-polymer iesh N={M + 1} P={P}: {rate:.3e} trajectory-steps/s")
    return res, sample


def spin_boson_extra(L):
    """Spin-boson Ehrenfest and FSSH: eps = Delta = 1, 100 modes of a Debye bath (eta = 0.25, wc = 5),
    Wigner-sampled at beta = 5, 1e5 trajectories to t = 20 / Delta with the population-correlation
    output every 50 steps; FP32 bath, FP64 electronics.  The step is 0.002 / Delta: the highest
    discretised mode (wc tan(pi/2 * 100/101) = 321) puts velocity Verlet beyond its stability limit
    at 0.01."""
    nm, wc, lam, beta = 100, 5.0, 0.625, 5.0
    j = np.arange(1, nm + 1)
    w = wc * np.tan(np.pi / 2 * (1 - j / (nm + 1)))
    c = np.sqrt(2 * lam / (nm + 1)) * w
    t = np.tanh(beta * w / 2)
    prm = np.concatenate([np.zeros(nm), np.sqrt(1 / (2 * w * t)), np.zeros(nm), np.sqrt(w / (2 * t))])
    ntraj, dt, nsteps = 100000, 0.002, 10000
    res = {}
    nsamp = 512  # first trajectories handed to the oracle check of the CPU-baseline leg
    sample = {"w": w, "c": c, "sdr": np.sqrt(1 / (2 * w * t)), "sdv": np.sqrt(w / (2 * t)), "n": nsamp, "dt": dt,
              "nsteps": nsteps, "save_every": 50, "nsave": nsteps // 50 + 1}
    for name, method in (("ehrenfest", L.METHOD_EHRENFEST), ("fssh", L.METHOD_FSSH)):
        E = L.Ensemble(0)
        E.set_problem(method, 2, 1, nm, np.ones(nm), dt, precision=L.PREC_FP32_NUCLEI)
        E.set_model(0, L.MODEL_SPIN_BOSON, np.concatenate([[1.0, 1.0], w, c]))
        E.sample_initial(L.SAMPLER_HARMONIC_WIGNER, ntraj, prm, KEY, initial_state=0, state_basis=1)
        st = E.run(0.0, nsteps * dt, KEY, save_every=50)
        E.reduce(L.OUT_POPULATION_CORRELATION, L.RED_MEAN, 4 * E.nsave)
        sample[name] = E.download(L.FIELD_SAVE_POP_DI)[:nsamp].copy()
        E.close()
        rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
        res[name] = {"ntraj": ntraj, "nsteps": nsteps, "dt": dt, "kernel_ms": st["kernel_ms"],
                     "traj_steps_per_s": rate, "grid": st["grid"], "block": st["block"]}
        _progress(f"spin-boson {name}: {rate:.3e} trajectory-steps/s")
    return res, sample


def ring_polymer_extra(L):
    """RPMD (one state) and ring-polymer surface hopping (two states) on the double well
    V = x^4/1200 - x^2/40 with a Gaussian-coupled second diabat, beta = 8, 1e5 ring polymers for 1000
    steps of dt = 1 at 8, 16, 32 and 64 beads; FP32 beads, FP64 electronics.  Initial conditions from the
    device path-integral Monte Carlo sampler (200 chain steps per ring polymer from x0 = -3.87, centroid
    moves of 3 bohr, the free ring polymer's mode widths without the second test of the last proposal:
    the discretised path-integral distribution).  Rates are ring-polymer steps/s (a ring polymer = one
    trajectory)."""
    mass, beta, ntraj, nsteps = 1836.0, 8.0, 100000, 1000
    pimc = [-3.87, 3.0, 200, 1.0, 1, 0]
    res = {"initial_conditions": "PIMC " + str(pimc)}
    sample = None
    for name, method, nstates, prm in (("rpmd", L.METHOD_RPMD, 1, [1 / 1200, 1 / 40, 0.0, 0.0, 1.0]),
                                       ("rpsh", L.METHOD_RPSH, 2, [1 / 1200, 1 / 40, 0.01, 0.005, 1.0])):
        for P in (8, 16, 32, 64):
            E = L.Ensemble(0)
            alg = L.ALG_BCB if method == L.METHOD_RPMD else L.ALG_BCB_TSIT5
            E.set_problem(method, nstates, 1, 1, [mass], 1.0, algorithm=alg, nbeads=P, precision=L.PREC_FP32_NUCLEI,
                          temperature=1.0 / beta)
            E.set_model(0, L.MODEL_DW_GAUSSIAN, prm)
            t0 = time.perf_counter()
            E.sample_initial(L.SAMPLER_RING_POLYMER_PIMC, ntraj, pimc, KEY)
            t_pimc = time.perf_counter() - t0
            keep = name == "rpsh" and P == 8
            if keep:  # the first ring polymers' initial beads for the oracle check of the CPU-baseline leg
                E.run(0.0, 0.0, KEY)
                n = 256
                sample = {"prm": prm, "mass": mass, "T": 1.0 / beta, "nsteps": nsteps, "save_every": 100,
                          "nsave": nsteps // 100 + 1, "r0": E.download(L.FIELD_R)[:n].copy(),
                          "v0": E.download(L.FIELD_V)[:n].copy()}
            st = E.run(0.0, float(nsteps), KEY, save_every=100)
            if keep:
                sample["state"] = E.download(L.FIELD_STATE)[:n].copy()
                sample["centroid"] = E.download(L.FIELD_SAVE_CENTROID)[:n].copy()
            E.close()
            rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
            res[f"{name}_P{P}"] = {"ntraj": ntraj, "nsteps": nsteps, "nbeads": P, "kernel_ms": st["kernel_ms"],
                                   "traj_steps_per_s": rate, "bead_steps_per_s": rate * P, "pimc_wall_s": t_pimc,
                                   "grid": st["grid"], "block": st["block"]}
            _progress(f"{name} P={P}: {rate:.3e} ring-polymer steps/s")
    return res, sample


# three Morse curves with one wall (D = 0.02, beta = 0.5, r = 4) offset by 0, 0.004 and 0.008, Gaussian-coupled:
# the mapping Hamiltonian weighs V_ii by (X_i^2 + P_i^2 - gamma) / 2, which can be negative, so walls of different
# steepness make its potential unbounded below (trajectories then run off); with one wall shape it stays bounded
MAP_MORSE3 = [0.02, 0.5, 4.0, 0.0, 0.02, 0.5, 4.0, 0.004, 0.02, 0.5, 4.0, 0.008,
              0.003, 0.5, 4.0, 0.002, 0.5, 4.2, 0.003, 0.5, 4.4]


def mapping_extra(L):
    """Spin mapping (SpinMappingW) integrated by MInt, FP64, 1e6 trajectories for 1000 steps with the
    diabatic populations and energies saved every 100 steps: Tully's model 1 (two states, mass 2000, x0 ~
    N(-6, 0.5), p0 ~ N(15, 1), dt = 1, occupied diabat 1) and three coupled Morse curves (MAP_MORSE3, mass
    20000, x0 ~ N(4.1, 0.1), v0 ~ N(0, 2e-4), dt = 2, occupied diabat 1).  The mapping variables are drawn on the
    device.  Rates are trajectory-steps/s of kernel time."""
    ntraj, nsteps, se = 1_000_000, 1000, 100
    res, sample = {}, {}
    cases = (("tully1_n2", L.MODEL_TULLY1, [0.01, 1.6, 0.005, 1.0], 2, 2000.0, [-6.0, 0.5, 15.0 / 2000.0, 1.0 / 2000.0], 1.0),
             ("morse_n3", L.MODEL_MORSE_N, [3] + MAP_MORSE3, 3, 20000.0, [4.1, 0.1, 0.0, 2e-4], 2.0))
    for name, kind, prm, n, mass, gauss, dt in cases:
        E = L.Ensemble(0)
        E.set_problem(L.METHOD_SPIN_MAPPING, n, 1, 1, [mass], dt, algorithm=L.ALG_MINT)
        E.set_model(0, kind, prm)
        E.sample_initial(L.SAMPLER_GAUSSIAN, ntraj, gauss, KEY, initial_state=0, state_basis=1)
        E.run(0.0, 0.0, KEY)
        m = 256   # first trajectories handed to the oracle check of the CPU-baseline leg
        sample[name] = {"kind": kind, "prm": prm, "n": n, "mass": mass, "dt": dt, "nsteps": nsteps, "save_every": se,
                        "r0": E.download(L.FIELD_R)[:m].copy(), "v0": E.download(L.FIELD_V)[:m].copy()}
        st = E.run(0.0, dt * nsteps, KEY, save_every=se)
        pop = E.reduce(L.OUT_DIABATIC_POPULATION, L.RED_MEAN, n * E.nsave).reshape(E.nsave, n)
        sample[name]["pop"] = E.download(L.FIELD_SAVE_POP_DI)[:m].copy()
        E.close()
        rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
        res[name] = {"ntraj": ntraj, "nsteps": nsteps, "dt": dt, "nstates": n, "kernel_ms": st["kernel_ms"],
                     "traj_steps_per_s": rate, "grid": st["grid"], "block": st["block"],
                     "final_diabatic_population": [float(x) for x in pop[-1]]}
        _progress(f"spin mapping {name}: {rate:.3e} trajectory-steps/s")
    # NRPMD by RingPolymerMInt: 1e5 ring polymers of 8 beads on Tully's model 1 (gamma 0.5, T = 1e-3, beads
    # ~ N(-5, 0.05), p0 ~ N(15, 1) shared by the beads, dt = 1), occupied diabat 1
    nrp, P = 100_000, 8
    rng = np.random.default_rng(12)
    r0 = -5.0 + 0.05 * rng.standard_normal((nrp, P))
    v0 = np.repeat((15.0 + rng.standard_normal((nrp, 1))) / 2000.0, P, axis=1)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_NRPMD, 2, 1, 1, [2000.0], 1.0, algorithm=L.ALG_RP_MINT, nbeads=P, temperature=1e-3,
                  mapping_gamma=0.5)
    E.set_model(0, L.MODEL_TULLY1, [0.01, 1.6, 0.005, 1.0])
    E.load_initial(r0, v0, None, state=np.zeros(nrp, np.int32))
    st = E.run(0.0, float(nsteps), KEY, save_every=se)
    pop = E.reduce(L.OUT_DIABATIC_POPULATION, L.RED_MEAN, 2 * E.nsave).reshape(E.nsave, 2)
    m = 128
    sample["nrpmd_P8"] = {"kind": 0, "prm": [0.01, 1.6, 0.005, 1.0], "n": 2, "P": P, "mass": 2000.0, "dt": 1.0,
                          "nsteps": nsteps, "save_every": se, "r0": r0[:m].copy(), "v0": v0[:m].copy(),
                          "pop": E.download(L.FIELD_SAVE_POP_DI)[:m].copy()}
    E.close()
    rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
    res["nrpmd_P8"] = {"ring_polymers": nrp, "nbeads": P, "nsteps": nsteps, "dt": 1.0, "kernel_ms": st["kernel_ms"],
                       "ring_polymer_steps_per_s": rate, "bead_steps_per_s": rate * P, "grid": st["grid"],
                       "block": st["block"], "final_diabatic_population": [float(x) for x in pop[-1]]}
    _progress(f"NRPMD P=8: {rate:.3e} ring-polymer steps/s")
    return res, sample


def _check_mapping(O, S, nthreads):
    """Spin mapping on the first trajectories of each workload against the FP64 oracle from the same initial
    conditions and mapping-variable draws: largest per-save diabatic population difference."""
    out = {}
    for name, d in S.items():
        n = d["n"]
        if "P" in d:   # NRPMD ring polymers
            k = len(d["r0"])
            m = O.make_model(O.MODEL_TULLY1, 2, 1, d["prm"])
            q0, p0 = O.rpmap_sample(O.RPMAP_NRPMD, n, d["P"], 0.5, k, KEY, state=np.zeros(k, np.int32))
            c = O.rpmap_run(O.RPMAP_NRPMD, m, n, d["P"], d["mass"], 1e-3, 0.5, d["r0"], d["v0"], q0, p0, dt=d["dt"],
                            nsteps=d["nsteps"], save_every=d["save_every"], nthreads=nthreads)
            out[name] = {"ring_polymers": k, "max_abs_err_population": float(np.abs(c["pop"] - d["pop"]).max())}
            continue
        m = (O.make_model(O.MODEL_TULLY1, 2, 1, d["prm"]) if d["kind"] == 0 else
             O.make_model(O.MODEL_MORSE_N, n, 1, d["prm"][1:]))
        k = len(d["r0"])
        X0, P0 = O.mapping_sample(n, np.zeros(k, np.int32), KEY)
        c = O.mapping_run(m, n, d["mass"], d["r0"][:, 0], d["v0"][:, 0], X0, P0, dt=d["dt"], nsteps=d["nsteps"],
                          save_every=d["save_every"], nthreads=nthreads)
        out[name] = {"trajectories": k, "max_abs_err_population": float(np.abs(c["pop"] - d["pop"]).max())}
    return out


# three-state Morse chain (the n-state Morse form, ThreeStateMorse-like values): [D, beta, r, c] per state,
# then [A, a, r] per pair (1,2), (1,3), (2,3)
MORSE3_FSSH = [0.02, 0.4, 4.0, 0.02, 0.02, 0.65, 4.5, 0.0, 0.003, 0.65, 6.0, 0.02,
               0.005, 32.0, 3.40, 0.0, 32.0, 4.8, 0.005, 32.0, 4.97]


def nstate_extra(L):
    """FSSH with three states (BABwithTsit5 on the three-state Morse chain MORSE3_FSSH, mass 20000, x0 ~ N(3, 0.15),
    kinetic energy U(0.02, 0.04), start on adiabatic state 1, 2000 steps of dt = 1, 1e5 trajectories) in FP32 nuclei
    and FP64, and one persistent launch mixing 5e4 two-state Tully-1 trajectories (x0 ~ N(-5, 0.5), p0 ~ N(60, 2)) with
    5e4 three-state Morse ones (one mass, 20000, for both).  Rates are trajectory-steps/s of kernel time."""
    ntraj, nsteps, mass = 100_000, 2000, 20000.0
    rng = np.random.default_rng(21)
    r0 = rng.normal(3.0, 0.15, size=(ntraj, 1))
    v0 = np.sqrt(2 * rng.uniform(0.02, 0.04, size=(ntraj, 1)) / mass)
    sig0 = np.zeros((ntraj, 3, 3), complex); sig0[:, 0, 0] = 1.0
    st0 = np.zeros(ntraj, np.int32)
    res, sample = {}, {}
    for name, prec in (("fssh_n3_fp32", L.PREC_FP32_NUCLEI), ("fssh_n3_fp64", L.PREC_FP64)):
        E = L.Ensemble(0)
        E.set_problem(L.METHOD_FSSH, 3, 1, 1, [mass], 1.0, precision=prec)
        E.set_model(0, L.MODEL_MORSE_N, [3] + MORSE3_FSSH)
        E.load_initial(r0, v0, sig0, state=st0)
        st = E.run(0.0, float(nsteps), KEY, save_every=100)
        pop = E.reduce(L.OUT_ADIABATIC_POPULATION, L.RED_MEAN, 3 * E.nsave).reshape(E.nsave, 3)
        if prec == L.PREC_FP64:
            m = 64
            sample = {"mass": mass, "nsteps": nsteps, "r0": r0[:m].copy(), "v0": v0[:m].copy(),
                      "state": E.download(L.FIELD_STATE)[:m].copy(), "r": E.download(L.FIELD_R)[:m].copy()}
        E.close()
        rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
        res[name] = {"ntraj": ntraj, "nsteps": nsteps, "kernel_ms": st["kernel_ms"], "traj_steps_per_s": rate,
                     "grid": st["grid"], "block": st["block"], "final_adiabatic_population": [float(x) for x in pop[-1]]}
        _progress(f"n-state {name}: {rate:.3e} trajectory-steps/s")
    # one launch mixing two- and three-state trajectories (model slot 0: Tully 1, slot 1: the Morse chain)
    half = ntraj // 2
    mid = np.concatenate([np.zeros(half, np.int32), np.ones(ntraj - half, np.int32)])
    rm = np.where(mid[:, None] == 0, rng.normal(-5.0, 0.5, size=(ntraj, 1)), r0)
    vm = np.where(mid[:, None] == 0, rng.normal(60.0, 2.0, size=(ntraj, 1)) / mass, v0)
    E = L.Ensemble(0)
    E.set_problem(L.METHOD_FSSH, 3, 1, 1, [mass], 1.0, precision=L.PREC_FP32_NUCLEI)
    E.set_model(0, L.MODEL_TULLY1, [0.01, 1.6, 0.005, 1.0])
    E.set_model(1, L.MODEL_MORSE_N, [3] + MORSE3_FSSH)
    E.load_initial(rm, vm, sig0, state=st0, model_id=mid)
    st = E.run(0.0, float(nsteps), KEY, term=(0, 10.0, L.TERM_ABS_GT))
    E.close()
    rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
    res["mixed_n2_n3_fp32"] = {"ntraj": ntraj, "nsteps_max": nsteps, "traj_steps": st["traj_steps"],
                               "kernel_ms": st["kernel_ms"], "traj_steps_per_s": rate}
    _progress(f"n-state mixed n2/n3: {rate:.3e} trajectory-steps/s")
    return res, sample


def _check_nstate(O, S, nthreads):
    """Three-state FSSH in FP64 on its first trajectories against the oracle's general-n path: final active
    states and positions."""
    k = len(S["r0"])
    m = O.make_model(O.MODEL_MORSE_N, 3, 1, MORSE3_FSSH)
    sig0 = np.zeros((k, 3, 3), complex); sig0[:, 0, 0] = 1.0
    c = O.run_ensemble(0, m, S["mass"], S["r0"], S["v0"], sig0, state0=np.zeros(k, np.int32), dt=1.0,
                       nsteps=S["nsteps"], key=KEY, nthreads=nthreads)
    return {"trajectories": k, "state_agreement": float(np.mean(c["state"] == S["state"])),
            "max_abs_err_r": float(np.abs(c["r"].ravel() - S["r"].ravel()).max())}


def cme_extra(L):
    """Classical master equation surface hopping on the one-mode Anderson-Holstein model (the reference's phonon
    relaxation test, test/Dynamics/cme.jl:57-88: hw = Gamma = 0.003, kT = 0.03, g = 0.005), 1e6 trajectories for
    1000 steps of dt = 5 (CME), and BCME (Gamma = 0.1, kT = 0.01, g = 0.015, bandwidth 100) on 1e6 trajectories for
    1000 steps of dt = 2; FP64.  Rates are trajectory-steps/s of kernel time."""
    res, sample = {}, {}
    cases = (("cme", L.METHOD_CME, 0.003, 0.003, 0.03, 0.005, 1_000_000, 5.0, 0.0),
             ("bcme", L.METHOD_BCME, 0.003, 0.1, 0.01, 0.015, 1_000_000, 2.0, 100.0))
    for name, method, hw, Gam, kT, g, ntraj, dt, W in cases:
        Ed = 0.0 if method == L.METHOD_CME else g ** 2 / hw
        prm = [hw, Ed, np.sqrt(2) * g, np.sqrt(Gam / (2 * np.pi))]
        samp = [0.0, np.sqrt(2 * kT / hw), 0.0, np.sqrt(2 * kT * hw)]
        E = L.Ensemble(0)
        E.set_problem(method, 2, 1, 1, [1 / hw], dt, algorithm=L.ALG_TSIT5, rescaling=L.RESCALE_OFF, temperature=kT,
                      bandwidth=W)
        E.set_model(0, L.MODEL_ANDERSON_HOLSTEIN_1, prm)
        E.sample_initial(L.SAMPLER_GAUSSIAN, ntraj, samp, KEY)
        st = E.run(0.0, 1000 * dt, KEY, save_every=100)
        ek = E.reduce(L.OUT_KINETIC_ENERGY, L.RED_MEAN, E.nsave)
        m = 256 if method == L.METHOD_CME else 128
        sample[name] = {"method": 0 if method == L.METHOD_CME else 1, "prm": prm, "samp": samp, "hw": hw, "kT": kT,
                        "dt": dt, "W": W, "state": E.download(L.FIELD_STATE)[:m].copy(),
                        "r": E.download(L.FIELD_R)[:m].copy(), "n": m}
        E.close()
        rate = st["traj_steps"] / (st["kernel_ms"] / 1e3)
        res[name] = {"ntraj": ntraj, "nsteps": 1000, "dt": dt, "kernel_ms": st["kernel_ms"], "traj_steps_per_s": rate,
                     "grid": st["grid"], "block": st["block"], "final_kinetic_energy": float(ek[-1])}
        _progress(f"{name}: {rate:.3e} trajectory-steps/s")
    return res, sample


def _check_cme(O, S, nthreads):
    """CME and BCME on their first trajectories against the oracle from the same draws: final states, and
    final positions where the states agree (the oracle's BCME integrals are adaptive Gauss-Kronrod quadrature,
    the device's a closed form)."""
    out = {}
    for name, C in S.items():
        n = C["n"]
        r0, v0 = O.sample_gaussian(n, 1, *C["samp"], KEY)
        m = O.make_model(O.MODEL_ANDERSON_HOLSTEIN_1, 2, 1, C["prm"])
        c = O.cme_run(C["method"], m, 1 / C["hw"], C["kT"], r0[:, 0], v0[:, 0], np.zeros(n, np.int32), dt=C["dt"],
                      nsteps=1000, bandwidth=C["W"] or 1.0, key=KEY, nthreads=nthreads)
        same = c["state"] == C["state"]
        out[name] = {"trajectories": n, "state_agreement": float(np.mean(same)),
                     "max_abs_err_r": float(np.abs(c["x"] - C["r"].ravel())[same].max()) if same.any() else None}
    return out


TRAIN_STEPS = 40000
TRAIN_BATCH = 2048


def surrogate_extra(L):
    """Learned form on the six-coordinate diatom-surface model: train the two-member BF16 surrogate
    on configurations the analytic FSSH visits, then FSSH on the surrogate vs on the analytic model
    (persistent kernel) and vs the plain port of the analytic model (FP64, one launch per step, state in
    HBM) on the same 1e6 trajectories, and the final-population errors."""
    import nqcd_mi355x as NQ
    analytic = NQ.DiatomSurface()
    mass = [12859.0, 12859.0]

    def inputs(ntraj, seed):
        rng = np.random.default_rng(seed)
        u = rng.normal(size=(ntraj, 3)); u /= np.linalg.norm(u, axis=1, keepdims=True)
        b = 2.2 + 0.1 * rng.normal(size=(ntraj, 1))
        c = np.concatenate([rng.normal(size=(ntraj, 2)), np.full((ntraj, 1), 6.0)], axis=1)
        r0 = np.concatenate([c + 0.5 * b * u, c - 0.5 * b * u], axis=1)
        v0 = 2e-4 * rng.normal(size=(ntraj, 6)); v0[:, 2] -= 3e-3; v0[:, 5] -= 3e-3
        return r0, v0

    def run(model, r0, v0, save_every=0, precision=L.PREC_FP32_NUCLEI, schedule=L.SCHED_PERSISTENT):
        E = L.Ensemble(0)
        E.set_problem(L.METHOD_FSSH, 2, 3, 2, mass, 2.0, precision=precision, schedule=schedule)
        E.set_model(0, model.kind, model.params())
        if model.kind == L.MODEL_MLP_SURROGATE:
            for k, p in enumerate(model.members):
                E.surrogate_set_params(k, p, model.norm)
        sig = np.zeros((len(r0), 2, 2), complex); sig[:, 0, 0] = 1
        E.load_initial(r0, v0, sig, state=np.zeros(len(r0), np.int32))
        st = E.run(0.0, 2.0 * 5000, KEY, save_every=save_every, term=(2, 8.0))
        out = {"stats": st, "final": E.reduce(L.OUT_FINAL_ADIABATIC_POPULATION, L.RED_MEAN, 2),
               "state": E.download(L.FIELD_STATE)}
        if save_every:
            out["configs"], out["steps"] = E.download(L.FIELD_SAVE_R), E.download(L.FIELD_STEPS)
        if model.kind == L.MODEL_MLP_SURROGATE:
            out["fallbacks"] = int(E.download(L.FIELD_FALLBACKS).sum())
        E.close()
        return out

    r0, v0 = inputs(4000, 41)
    vis = run(analytic, r0, v0, save_every=20)
    keep = np.arange(vis["configs"].shape[1])[None, :] * 20 <= vis["steps"][:, None]
    X = vis["configs"][keep]
    t0 = time.perf_counter()
    model, hist = NQ.train_surrogate(analytic, X, steps=TRAIN_STEPS, batch=TRAIN_BATCH, lr=3e-3, lr_final=1e-5, threshold=1e-3)
    t_train = time.perf_counter() - t0
    _progress(f"surrogate trained on {len(X)} configurations in {t_train:.1f} s")
    r1, v1 = inputs(1000000, 43)
    ref = run(analytic, r1, v1)
    sur = run(model, r1, v1)
    plain = run(analytic, r1, v1, precision=L.PREC_FP64, schedule=L.SCHED_STEP_PER_LAUNCH)
    n = 2000  # the first trajectories for the oracle check of the CPU-baseline leg
    sample = {"r0": r1[:n].copy(), "v0": v1[:n].copy(), "params": analytic.params(), "mass": np.array(mass),
              "nsteps": 5000, "term": 8.0, "surrogate_state": sur["state"][:n], "analytic_state": ref["state"][:n]}
    rate = lambda o: o["stats"]["traj_steps"] / (o["stats"]["kernel_ms"] / 1e3)
    p_ref, p_sur, p_plain = float(ref["final"][1]), float(sur["final"][1]), float(plain["final"][1])
    plain_rate = plain["stats"]["traj_steps"] / (plain["stats"]["wall_ms"] / 1e3)
    _progress(f"surrogate FSSH {rate(sur):.3e} / analytic {rate(ref):.3e} / plain port {plain_rate:.3e} "
              f"trajectory-steps/s")
    # BF16 MFMA work of one trajectory-step: two members x (forward 128x128 and 16(3)x128 products +
    # three reverse-mode 128x128 products for the input gradient)
    flop = 2 * (4 * 2 * 128 * 128 + 2 * 16 * 128)
    return {"workload": "diatom-surface FSSH, 6 coordinates, 1e6 trajectories, dt 2, t_max 1e4, terminate z1 > 8",
            "training_configurations": int(len(X)), "training_steps": TRAIN_STEPS, "training_batch": TRAIN_BATCH, "training_s": t_train,
            "final_loss": [float(hist[0, -100:].mean()), float(hist[1, -100:].mean())],
            "surrogate_traj_steps_per_s": rate(sur), "analytic_traj_steps_per_s": rate(ref),
            "surrogate_kernel_ms": sur["stats"]["kernel_ms"], "analytic_kernel_ms": ref["stats"]["kernel_ms"],
            "plain_traj_steps_per_s": plain_rate, "plain_wall_ms": plain["stats"]["wall_ms"],
            "plain_launches": plain["stats"]["launches"],
            "learned_vs_plain": rate(sur) / plain_rate, "batched_vs_plain": rate(ref) / plain_rate,
            "final_excited_population": {"analytic": p_ref, "surrogate": p_sur, "plain_fp64": p_plain},
            "pop_error_vs_analytic": abs(p_sur - p_ref), "pop_error_vs_plain_fp64": abs(p_sur - p_plain),
            "pop_sigma": float(np.sqrt(p_ref * (1 - p_ref) / len(r1))),
            "fallback_fraction": sur["fallbacks"] / sur["stats"]["traj_steps"],
            "bf16_flop_per_traj_step": flop,
            "achieved_bf16_TFLOPs": flop * rate(sur) / 1e12, "frac_bf16_peak": flop * rate(sur) / 1e12 / 2500.0}, sample


TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "fssh2_persistent_c2_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the headline kernel from the committed counter summary (the line
    `traffic_bytes_per_launch: N`), or None when absent."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-port comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip the IESH and surrogate measurements")
    ap.add_argument("--only", choices=["iesh", "spin_boson", "ring_polymer", "learned", "mapping", "nstate", "cme"],
                    help="run one of the other workloads alone and print its JSON object (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB three times (the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 1 << 30, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    from nqcd_mi355x import _lib as L
    if args.only:
        fn = {"iesh": iesh_extra, "spin_boson": spin_boson_extra, "ring_polymer": ring_polymer_extra,
              "learned": surrogate_extra, "mapping": mapping_extra, "nstate": nstate_extra,
              "cme": cme_extra}[args.only]
        print(json.dumps({args.only: fn(L)[0]}))
        return

    # --- batched: persistent work-queue kernel, FP32 nuclei / FP64 electronics -------------------
    E = _ensemble(L, L.PREC_FP32_NUCLEI, L.SCHED_PERSISTENT)
    # launch shape of the persistent kernel chosen by the library's autotuner on whole passes with the
    # termination predicate (untimed), so the choice includes the tail of uneven trajectory lengths
    tuning = E.autotune(0.0, TMAX, KEY, term=(0, XTERM))
    _progress(f"autotune: block {tuning['block']}, workgroups/CU {tuning['workgroups_per_cu'] or 'max'} "
              f"({tuning['best_ms']:.2f} ms vs default {tuning['default_ms']:.2f} ms on the sample)")
    for _ in range(args.warmup):
        st = E.run(0.0, TMAX, KEY, term=(0, XTERM))
        _progress(f"warmup pass: {st['kernel_ms']:.1f} ms, {st['traj_steps']} trajectory-steps")
    kernel_ms, wall_ms, steps_total = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = E.run(0.0, TMAX, KEY, term=(0, XTERM))
        kernel_ms.append(st["kernel_ms"]); wall_ms.append(st["wall_ms"]); steps_total.append(st["traj_steps"])
        _progress(f"timed pass: {st['kernel_ms']:.1f} ms kernel, {st['wall_ms']:.1f} ms wall")
    t_timed = time.perf_counter() - t0
    assert len(set(steps_total)) == 1, "runs must be identical"
    traj_steps = steps_total[0]
    ms_per_step = 1e3 * t_timed / args.steps
    value = traj_steps / (t_timed / args.steps)
    avg_kernel_ms = float(np.mean(kernel_ms))
    pops32 = _final_pops(L, E)
    states32 = E.download(L.FIELD_STATE)
    grid, block = st["grid"], st["block"]
    E.close()

    # --- FP64 persistent (the parity path) on the same ensemble ----------------------------------
    E64 = _ensemble(L, L.PREC_FP64, L.SCHED_PERSISTENT)
    st64 = E64.run(0.0, TMAX, KEY, term=(0, XTERM))
    pops64 = _final_pops(L, E64)
    fp64_rate = st64["traj_steps"] / (st64["kernel_ms"] / 1e3)
    _progress(f"fp64 persistent pass: {st64['kernel_ms']:.1f} ms")
    E64.close()
    pop_err = float(np.abs(pops32 - pops64).max())
    p = pops64[:2]
    sigma = float(np.sqrt(np.max(p * (1 - p)) / NTRAJ))

    # --- plain GPU port: one launch per step, HBM state, dense FP64 -------------------------------
    plain = None
    if not args.no_plain:
        Ep = _ensemble(L, L.PREC_FP64, L.SCHED_STEP_PER_LAUNCH)
        stp = Ep.run(0.0, TMAX, KEY, term=(0, XTERM))
        popsp = _final_pops(L, Ep)
        _progress(f"plain pass: {stp['wall_ms']:.1f} ms wall, {stp['launches']} launches")
        Ep.close()
        plain = {"value": stp["traj_steps"] / (stp["wall_ms"] / 1e3), "kernel_ms": stp["kernel_ms"],
                 "wall_ms": stp["wall_ms"], "launches": stp["launches"],
                 "pop_diff_vs_fp64_persistent": float(np.abs(popsp - pops64).max())}

    achieved = FLOP_PER_STEP * traj_steps / (avg_kernel_ms / 1e3) / 1e12
    algo_bytes = NTRAJ * (8 + 8 + 32 + 4 + 4 + 4 + 8 + 8 + 32 + 12)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "trajectory-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 nuclei / f64 electronics",
        "data": "synthetic (device-sampled Philox initial conditions)",
        "config": {"workload": "batched FSSH, 1e6 trajectories mixing Tully 1/2/3, p0 grid 5..35 a.u., "
                               "dt 1 a.u., t_max 10000 a.u., terminate |x|>10",
                   "ntraj": NTRAJ, "traj_steps_per_pass": int(traj_steps), "parallelism": "1 GPU",
                   "schedule": "persistent work-queue kernel", "grid": grid, "block": block},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this workload, scaled by a 1 GiB copy's "
                                       "reading (profiles/fssh2_persistent_c2_traffic.md), not measured in this run",
                     "kernel": "fssh2_persistent<1,float>", "kernel_avg_ms": avg_kernel_ms,
                     "flop_per_traj_step": FLOP_PER_STEP,
                     "hbm_algorithmic_GBs": algo_bytes / (avg_kernel_ms / 1e3) / 1e9},
        "autotune": tuning,
        "pop_error_vs_fp64": pop_err,
        "pop_sigma": sigma,
        "fp64_persistent": {"value": fp64_rate, "kernel_ms": st64["kernel_ms"]},
        "plain_gpu": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
    }
    samples = {}
    if not args.no_extras:
        out["iesh"], samples["iesh"] = iesh_extra(L)
        out["spin_boson"], samples["spin_boson"] = spin_boson_extra(L)
        out["ring_polymer"], samples["ring_polymer"] = ring_polymer_extra(L)
        out["learned"], samples["learned"] = surrogate_extra(L)
        out["mapping"], samples["mapping"] = mapping_extra(L)
        out["nstate"], samples["nstate"] = nstate_extra(L)
        out["cme"], samples["cme"] = cme_extra(L)
    if not args.no_cpu:
        nthreads = min(CPU_THREADS, os.cpu_count() or 1)
        _progress(f"cpu baseline on {nthreads} threads and on one core")
        out["cpu_baseline"] = cpu_baseline(L, states32, nthreads, samples)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
