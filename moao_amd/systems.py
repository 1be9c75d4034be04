"""Seeded synthetic MOAO systems written in the reference's parameter-file formats.

Each generator writes ``sys-params.txt`` and ``prof<k>-atmos-night0.txt`` files that the
reference (SysParams::read / AtmParams::read) and this library read identically, so both
paths see the same bits.  Field order: common/tomo_struct/sysParams.cpp:165-207 and
atmParams.cpp:93-101.

Workloads (sizes follow the valid-subaperture rule of tomo_struct.cpp:403-445):
  small   D=8 m, obs 0.28, 3 LGS (90 km) on a 60" ring + on-axis truth, 10x10, 2 layers
          (0, 10 km), r0=0.15, L0=25                                     -> n=456, m=152
  elt     D=39 m, obs 0.28, 6 LGS (90 km) on a 7.5' ring + 2 NGS within 5' + truth, 40x40,
          10 layers U(0, 20 km), r0=0.129, L0=25                        -> n=18688, m=2336
  elt74   D=39 m, obs 0.28, 3 LGS on the 7.5' ring + 2 NGS + truth, 74x74, 10 layers
                                                                         -> n=39720, m=7944
  custom  any (nlgs, nngs, nssp, nlayer, diam, r0) for the request stream
Ring sizes are read as diameters, as the reference's synthetic SysParams constructor does with
its 7.4' LGS ring (sysParams.cpp:94-97).  Cn2 fractions are Dirichlet(1,...,1) from the seed
and cn2[l] = f_l * r0^(-5/3) (the reference derives r0 from sum(cn2), tomo_struct.cpp:317-319).
Fields the workloads leave open take the values of the reference's check fixtures
(datafile/check/*/sys-params.txt): RON 3, LGS flux 6e7, spot 1", depth 5 km, lgs_cst 1,
NGS magnitude 13, pixel sizes 0.3"/1", transmissions 0.425/0.382/0.84.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class SystemSpec:
    diam: float = 39.0
    obs: float = 0.28
    nssp: int = 40
    nlgs: int = 6
    nngs: int = 2
    lgs_ring_diam_as: float = 450.0       # 7.5 arcmin
    ngs_field_diam_as: float = 300.0      # within 5 arcmin
    lgs_alt: float = 90000.0
    nlayer: int = 10
    r0: float = 0.129
    L0: float = 25.0
    h_max: float = 20000.0
    fixed_h: List[float] = field(default_factory=list)
    lgs_cst: float = 1.0
    seed: int = 1234
    ntarget: int = 1                      # truth directions (the first on axis)
    target_field_diam_as: float = 300.0   # the others uniform in the disk of this field


WORKLOADS = {
    "small": SystemSpec(diam=8.0, nssp=10, nlgs=3, nngs=0, lgs_ring_diam_as=60.0, nlayer=2,
                        r0=0.15, fixed_h=[0.0, 10000.0]),
    "elt": SystemSpec(diam=39.0, nssp=40, nlgs=6, nngs=2, nlayer=10, r0=0.129),
    "elt74": SystemSpec(diam=39.0, nssp=74, nlgs=3, nngs=2, nlayer=10, r0=0.129),
}


def _fmt(v):
    if isinstance(v, (list, tuple, np.ndarray)):
        return " ".join(repr(float(x)) if not isinstance(x, (int, np.integer)) else str(int(x))
                        for x in v)
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_sys(path: str, spec: SystemSpec, rng: np.random.Generator):
    nW = spec.nlgs + spec.nngs + 1
    ax, ay = [], []
    rad = spec.lgs_ring_diam_as / 2.0
    for k in range(spec.nlgs):
        a = 2.0 * math.pi * k / max(spec.nlgs, 1)
        ax.append(rad * math.cos(a))
        ay.append(rad * math.sin(a))
    for _ in range(spec.nngs):
        # uniform in the disk of the NGS field
        r = spec.ngs_field_diam_as / 2.0 * math.sqrt(rng.uniform())
        a = rng.uniform(0, 2 * math.pi)
        ax.append(r * math.cos(a))
        ay.append(r * math.sin(a))
    ax.append(0.0)
    ay.append(0.0)
    tx, ty = [0.0], [0.0]
    for _ in range(spec.ntarget - 1):
        r = spec.target_field_diam_as / 2.0 * math.sqrt(rng.uniform())
        a = rng.uniform(0, 2 * math.pi)
        tx.append(r * math.cos(a))
        ty.append(r * math.sin(a))
    gs_alt = [1.0 / spec.lgs_alt] * spec.nlgs + [0.0] * (spec.nngs + 1)
    gs_type = [2] * spec.nlgs + [1] * (spec.nngs + 1)
    rows = [
        ("tel diam in meters", spec.diam), ("cent obs", spec.obs), ("Tframe s", 0.004),
        ("N WFS", nW), ("nlgs", spec.nlgs), ("nTarget", spec.ntarget),
        ("Nssp n subaps per wfs", spec.nssp),
        ("GsAlt", gs_alt), ("type", gs_type), ("alphaX in arcsec", ax), ("alphaY in arcsec", ay),
        ("XPup", [0.0] * nW), ("YPup", [0.0] * nW), ("thetaML", [0.0] * nW),
        ("thetaCam", [0.0] * nW), ("sensibilite", [1.0] * nW), ("tracking", [0.0] * 3),
        ("pasDphi", 0.0001), ("ncpu", 1), ("magnitude of NGS", [13.0] * (nW - spec.nlgs)),
        ("Flux for LGS", 6e7), ("pixel size for ngs", 0.3), ("pixel size for lgs", 1.0),
        ("wave length for NGS (meters)", 650e-9), ("wave length for LGS (meters)", 589e-9),
        ("bandwidth", 3300e-10), ("transmission for NGS", 0.425),
        ("transmission for LGS", 0.382), ("atmosphere transmission", 0.84),
        ("Read Out Noise (nb of e-)", 3), ("lgs_cst", spec.lgs_cst), ("spot_width arcsec", 1.0),
        ("lgs_alt (m)", spec.lgs_alt), ("lgs_depth (m)", 5000.0),
        ("targetX  in arcsec", tx), ("targetY  in arcsec", ty),
    ]
    with open(path, "w") as fh:
        for name, v in rows:
            fh.write(f"{name}\n{_fmt(v)}\n")


def write_atm(path: str, spec: SystemSpec, rng: np.random.Generator, r0: float | None = None):
    L = spec.nlayer
    r0 = spec.r0 if r0 is None else r0
    frac = rng.dirichlet(np.ones(L))
    cn2 = frac * r0 ** (-5.0 / 3.0)
    if spec.fixed_h:
        h = list(spec.fixed_h[:L])
    else:
        h = sorted(rng.uniform(0.0, spec.h_max, size=L).tolist())
    with open(path, "w") as fh:
        fh.write(f"Nlayer\n{L}\nr0 @ wfs lambda\n{_fmt(r0)}\ncn2 ESO units\n{_fmt(cn2)}\n"
                 f"h in meters\n{_fmt(h)}\nl0 in meters\n{_fmt([spec.L0] * L)}\n")


def make_system(outdir: str, spec: SystemSpec | str, nprof: int = 1):
    """Writes sys-params.txt and profiles prof-1 .. prof{nprof-2}; returns (sys, [atm...])."""
    if isinstance(spec, str):
        spec = WORKLOADS[spec]
    os.makedirs(outdir, exist_ok=True)
    rng = np.random.default_rng(spec.seed)
    sys_path = os.path.join(outdir, "sys-params.txt")
    write_sys(sys_path, spec, rng)
    atms = []
    for k in range(-1, nprof - 1):
        p = os.path.join(outdir, f"prof{k}-atmos-night0.txt")
        write_atm(p, spec, rng)
        atms.append(p)
    return sys_path, atms


def valid_subaps(nssp: int, obs: float, diam: float = 39.0) -> int:
    """Count of the annulus rule (tomo_struct.cpp:403-445) for one WFS."""
    pas = diam / nssp
    lo = -diam / 2.0
    rtel2 = diam * diam / 4.0
    if nssp != 7 or (obs <= 0.285 or obs >= 0.29):
        robs2 = diam * obs / 2.0 * diam * obs / 2.0
    else:
        robs2 = diam * 0.285 / 2.0 * diam * 0.285 / 2.0
    c = lo + pas / 2.0 * (2.0 * np.arange(nssp) + 1.0)
    r2 = c[None, :] ** 2 + c[:, None] ** 2
    return int(np.count_nonzero(~((r2 < robs2) | (r2 >= rtel2))))


def dx_matrix(nact: int, nmeasts: int, nmeas: int, seed: int = 7) -> np.ndarray:
    """Command matrix Dx (nact x nmeasts) ~ N(0, 1/nmeas), FP64, column-major."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.normal(0.0, 1.0 / math.sqrt(nmeas), size=(nact, nmeasts)))


def request_stream(outdir: str, nreq: int, seed: int = 2024, n_max: int = 40000):
    """Systems of a serving stream: each request draws n_wfs ~ U{4..9} measuring sensors (two
    thirds LGS, the rest NGS, plus the truth sensor) and Nssp ~ U{20..74} on the D = 39 m
    telescope, redrawn while n = 2 n_wfs Nvalid(Nssp) > n_max, 10 layers with fresh seeded Cn2
    fractions and r0 ~ U(0.08, 0.2) m.  Returns [(sys_file, atm_file, n, m)], one directory per
    request under outdir."""
    rng = np.random.default_rng(seed)
    reqs = []
    while len(reqs) < nreq:
        nw = int(rng.integers(4, 10))
        nssp = int(rng.integers(20, 75))
        nv = valid_subaps(nssp, 0.28)
        n = 2 * nw * nv
        if n > n_max:
            continue
        nlgs = max(1, (2 * nw + 2) // 3)
        spec = SystemSpec(diam=39.0, nssp=nssp, nlgs=nlgs, nngs=nw - nlgs, nlayer=10,
                          r0=float(rng.uniform(0.08, 0.2)), seed=int(rng.integers(1 << 30)))
        d = os.path.join(outdir, f"r{len(reqs)}")
        os.makedirs(d, exist_ok=True)
        s, atms = make_system(d, spec)
        reqs.append((s, atms[0], n, 2 * nv))
    return reqs
