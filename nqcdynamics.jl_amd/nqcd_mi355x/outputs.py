"""Ensemble outputs (src/DynamicsOutputs.jl, src/TimeCorrelationFunctions.jl) and
reductions (src/Ensembles/reductions.jl).

Each output turns the device results of one run into the per-trajectory value the
reference's `f(sol, i)` returns.  Mean/Sum reductions of the outputs that have a
device reduction run on the GPU in a fixed order; the others reduce on the host.
"""
import os

import numpy as np

from . import _lib as L
from .distributions import Adiabatic, Diabatic


class Output:
    name = None
    device_reduction = None   # nqcd_output_kind with a device-side Mean/Sum

    def per_trajectory(self, res):
        raise NotImplementedError


class _FnOutput(Output):
    def __init__(self, name, fn, device_reduction=None):
        self.name, self._fn, self.device_reduction = name, fn, device_reduction

    def per_trajectory(self, res):
        return self._fn(res)

    def __repr__(self):
        return self.name


# OutputAdiabaticPopulation / OutputDiabaticPopulation (DynamicsOutputs.jl:186-203): one n-vector per save.
OutputAdiabaticPopulation = _FnOutput("OutputAdiabaticPopulation", lambda r: list(r["pop_ad"]), L.OUT_ADIABATIC_POPULATION)
OutputDiabaticPopulation = _FnOutput("OutputDiabaticPopulation", lambda r: list(r["pop_di"]), L.OUT_DIABATIC_POPULATION)
# OutputDiscreteState (DynamicsOutputs.jl:178): active state (1-based) per save.
OutputDiscreteState = _FnOutput("OutputDiscreteState", lambda r: list(r["save_state"] + 1))
# OutputSurfaceHops (DynamicsOutputs.jl:235-243): accepted hops.
OutputSurfaceHops = _FnOutput("OutputSurfaceHops", lambda r: r["nhops"])
# OutputFinalTime (DynamicsOutputs.jl:228-229).
OutputFinalTime = _FnOutput("OutputFinalTime", lambda r: r["t0"] + r["steps"] * r["dt"])
# OutputCentroidPosition (DynamicsOutputs.jl:73-79): ring-polymer centroid per save.
OutputCentroidPosition = _FnOutput("OutputCentroidPosition", lambda r: list(r["centroid"]), L.OUT_CENTROID_POSITION)
# OutputKineticEnergy / OutputTotalEnergy (DynamicsOutputs.jl:90-98): per save, two-state one-dof
# FSSH / Ehrenfest (potential: fssh.jl:150-154, ehrenfest.jl:85-95).
OutputKineticEnergy = _FnOutput("OutputKineticEnergy", lambda r: list(r["ekin"]), L.OUT_KINETIC_ENERGY)
OutputTotalEnergy = _FnOutput("OutputTotalEnergy", lambda r: list(r["ekin"] + r["epot"]), L.OUT_TOTAL_ENERGY)
# OutputPosition / OutputVelocity (DynamicsOutputs.jl:62-66, 35-39): per save (CME / BCME, classical).
OutputPosition = _FnOutput("OutputPosition", lambda r: list(r["save_r"]), L.OUT_POSITION)
OutputVelocity = _FnOutput("OutputVelocity", lambda r: list(r["save_v"]), L.OUT_VELOCITY)
# OutputFinal (DynamicsOutputs.jl:222): final dynamics variables.
OutputFinal = _FnOutput("OutputFinal", lambda r: {"v": r["v"], "r": r["r"], **(
    {"sigma": r["sigma"], "state": r["state"] + 1} if r.get("sigma") is not None else
    {"qmap": r["qmap"], "pmap": r["pmap"]} if "qmap" in r else
    {"state": r["state"] + 1} if "state" in r else {})})


class OutputStateResolvedScattering1D(Output):
    """Final populations split by the sign of the first coordinate (DynamicsOutputs.jl:313-338)."""
    name = "OutputStateResolvedScattering1D"

    def __init__(self, sim, kind="adiabatic"):
        if kind not in ("adiabatic", "diabatic"):
            raise ValueError(f"{kind} not recognised. Only `:diabatic` or `:adiabatic` accepted.")
        self.sim, self.kind = sim, kind
        self.device_reduction = L.OUT_SCATTERING_1D if kind == "adiabatic" else L.OUT_SCATTERING_1D_DIABATIC

    def per_trajectory(self, res):
        if self.kind == "diabatic" and "final_pop_di" not in res:
            raise ValueError("diabatic final populations exist for two-state one-dof FSSH / Ehrenfest")
        pops = res["final_pop_ad"] if self.kind == "adiabatic" else res["final_pop_di"]
        x = res["r"].reshape(-1)[0]
        z = np.zeros_like(pops)
        return {"reflection": z if x > 0 else pops, "transmission": pops if x > 0 else z}


class PopulationCorrelationFunction(Output):
    """norm K(0) P(t)^T per save point (TimeCorrelationFunctions.jl:26-40,61-88): K the initial estimator (the
    population itself but for eCMM's mapping kernel), norm 1 (F for eCMM, cmm.jl:114-119)."""
    name = "PopulationCorrelationFunction"

    def __init__(self, sim, statetype=None):
        self.sim = sim
        self.statetype = statetype if statetype is not None else Diabatic()
        diab = isinstance(self.statetype, Diabatic) or self.statetype is Diabatic
        self.diabatic = diab
        self.device_reduction = L.OUT_POPULATION_CORRELATION if diab else None

    def per_trajectory(self, res):
        P = res["pop_di"] if self.diabatic else res["pop_ad"]
        K0 = res.get("pop_init", P[0]) if self.diabatic else P[0]
        norm = res.get("corr_norm", 1.0)
        return [norm * np.outer(K0, p) for p in P]


# --- reductions -----------------------------------------------------------------------------
class SortByTrajectoryReduction:
    """One dictionary per trajectory (reductions.jl:54-55)."""


class SortByOutputReduction:
    """One list per output name (reductions.jl:117-143)."""


class MeanReduction:
    """Mean over trajectories (reductions.jl:33-52)."""


class SumReduction:
    """Sum over trajectories (reductions.jl:5-31)."""


class FileReduction:
    """Writes each trajectory's output dictionary to an HDF5 file (reductions.jl:57-91): group
    `trajectory_<id>` (1-based trajectory ids), one dataset per key, in HDF5.jl's layout (see h5.py).
    A filename without a .h5 / .hdf5 extension gets .h5 in place of its extension."""

    def __init__(self, filename):
        root, ext = os.path.splitext(str(filename))
        self.filename = str(filename) if ext in (".h5", ".hdf5") else root + ".h5"

    def __call__(self, u, batch, ids):
        from . import h5
        with h5.File(self.filename) as f:
            for i, traj in zip(ids, batch):
                with f.create_group(f"trajectory_{i}") as g:
                    for k, value in traj.items():
                        g.write(str(k), h5.file_order(value))
        return f"Output written to {self.filename}.", False


def output_name(f):
    return f.name if isinstance(f, Output) else getattr(f, "__name__", type(f).__name__)
