"""SimParams of the reference as a Python mirror, plus the workload presets.

SimParams::Reset defaults come from simulation/parameters_sim.cpp:7-55; the
derived constants (ComputeLame, ComputeHelperVariables) from
parameters_sim.cpp:122-140.  Strength parameters that the workloads do not
give take the values of input/n2k/nares2k.json:22-31.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, fields

from . import abi


@dataclass
class SimParams:
    GridXTotal: int = 256
    GridYTotal: int = 128
    cellsize: float = 1.0
    InitialTimeStep: float = 3.0e-5
    AnimationFramePeriod: float = 200.0
    IceDensity: float = 916.0
    PoissonsRatio: float = 0.3
    YoungsModulus: float = 5.0e8
    IceCompressiveStrength: float = 100e6
    IceTensileStrength: float = 10e6
    IceShearStrength: float = 1e6
    IceTensileStrength2: float = 10e6
    DP_phi: float = 62.0
    DP_threshold_p: float = 1e4
    RidgeFormationCoeff: float = 30.0
    ParticleVolume: float = 0.25
    sea_water_density: float = 1030.0
    waterDragEffectiveLinear: float = 0.01
    waterDragEffectiveQuadratic: float = 0.01
    windStressToAccel: float = 1.0 / 916.0
    tpb_P2G: int = 256
    tpb_Upd: int = 512
    tpb_G2P: int = 128
    flags: int = 0

    # derived (ComputeLame / ComputeHelperVariables)
    def lame(self):
        E, nu = self.YoungsModulus, self.PoissonsRatio
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
        mu = E / (2 * (1 + nu))
        kappa = mu * 2.0 / 3.0 + lam
        return lam, mu, kappa

    @property
    def UpdateEveryNthStep(self) -> int:
        return int(self.AnimationFramePeriod / self.InitialTimeStep)

    @property
    def ParticleMass(self) -> float:
        return self.ParticleVolume * self.IceDensity

    def to_c(self) -> abi.IceParams:
        p = abi.IceParams()
        for f in fields(self):
            setattr(p, f.name, getattr(self, f.name))
        lam, mu, kappa = self.lame()
        p.lambda_, p.mu, p.kappa = lam, mu, kappa
        p.ParticleMass = self.ParticleMass
        p.UpdateEveryNthStep = self.UpdateEveryNthStep
        p.cellsize_inv = 1.0 / self.cellsize
        p.Dp_inv = 4.0 / self.cellsize ** 2
        p.dt_vol_Dpinv = self.InitialTimeStep * self.ParticleVolume * p.Dp_inv
        p.vmax = 0.25 * self.cellsize / self.InitialTimeStep
        return p

    def replace(self, **kw) -> "SimParams":
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update(kw)
        return SimParams(**d)


def sea_ice_params(gx: int, gy: int, *, dt: float = 1e-4, cohesion: float = 1e5, phi_deg: float = 30.0,
                   drag_cd: float = 0.005, ice_thickness: float = 1.0, frame_steps: int = 100,
                   flags: int = 0, cellsize: float = 1.0) -> SimParams:
    """Material and forcing of the MPM workloads: E = 1e9 Pa, nu = 0.3,
    rho = 916, Drucker-Prager phi with cohesion c (DP_threshold_p = -c/tan phi,
    the apex of q = (p - DP_threshold_p) tan phi used at kernels.cuh:422),
    quadratic water drag rho_w Cd / (rho_i h_ice), wind stress acceleration
    1/(rho_i h_ice), cell size h (m).  ParticleVolume is set by the scene
    generator."""
    rho_i, rho_w = 916.0, 1030.0
    return SimParams(
        GridXTotal=gx, GridYTotal=gy, cellsize=cellsize, InitialTimeStep=dt,
        AnimationFramePeriod=frame_steps * dt * (1 + 1e-9),
        IceDensity=rho_i, PoissonsRatio=0.3, YoungsModulus=1e9,
        IceCompressiveStrength=100e6, IceTensileStrength=3e6, IceShearStrength=2.5e6,
        IceTensileStrength2=3e6, DP_phi=phi_deg,
        DP_threshold_p=-cohesion / math.tan(math.radians(phi_deg)),
        RidgeFormationCoeff=0.5, sea_water_density=rho_w,
        waterDragEffectiveLinear=0.0,
        waterDragEffectiveQuadratic=rho_w * drag_cd / (rho_i * ice_thickness),
        windStressToAccel=1.0 / (rho_i * ice_thickness),
        flags=flags,
    )
