"""Reader/writer for the reference's parameter text files (oracle side; test infrastructure).

Format (common/io/IO.cpp.in:64-78, common/io/named.cpp:19-31): every field is one comment line
followed by one value line, in a fixed order; the comment text is not checked.  Vectors are the
whitespace-separated numbers on the value line, up to the first token that is not a number
(common/io/IO.cpp.in:35-44 uses istream_iterator, so "25 25 same" reads [25, 25]).

Field order: SysParams::getParams (common/tomo_struct/sysParams.cpp:165-207) and
AtmParams::getParams (common/tomo_struct/atmParams.cpp:93-101).  nNgs = nW - nLgs - 1
(sysParams.cpp:146).

Pinned by: tests/test_oracle_tomo.py compares every Tomo field built from these readers with
the golden vectors written by the hand-built reference for datafile/check/nx456_* fixtures.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

# (name, kind) with kind in {"f": float scalar, "i": int scalar, "vf": float vector, "vi": int vector}
SYS_FIELDS = [
    ("diam", "f"), ("obs", "f"), ("tFrame", "f"), ("nW", "i"), ("nLgs", "i"), ("nTarget", "i"),
    ("nssp", "i"), ("gsAlt", "vf"), ("gsType", "vi"), ("alphaX_as", "vf"), ("alphaY_as", "vf"),
    ("XPup", "vf"), ("YPup", "vf"), ("thetaML", "vf"), ("thetaCam", "vf"), ("sensibilite", "vf"),
    ("tracking", "vf"), ("pasDphi", "f"), ("ncpu", "i"), ("mrNGS", "vf"), ("lgsFlux", "f"),
    ("ngsPixSize", "f"), ("lgsPixSize", "f"), ("lambdaNGS", "f"), ("lambdaLGS", "f"),
    ("bdw_m", "f"), ("throughNGS", "f"), ("throughLGS", "f"), ("throughAtm", "f"), ("RON", "i"),
    ("lgsCst", "f"), ("spotWidth", "f"), ("lgsAlt", "f"), ("lgsDepth", "f"),
    ("targetX_as", "vf"), ("targetY_as", "vf"),
]
ATM_FIELDS = [("nLayer", "i"), ("r0", "f"), ("cn2", "vf"), ("h", "vf"), ("L0", "vf")]


def _numbers(line: str, conv):
    out = []
    for tok in line.split():
        try:
            out.append(conv(tok))
        except ValueError:
            break
    return out


def _read(path: str, fields):
    with open(path) as fh:
        lines = fh.read().split("\n")
    vals = {}
    pos = 0
    for name, kind in fields:
        pos += 1  # comment line
        if pos >= len(lines):
            raise ValueError(f"{path}: truncated before field {name}")
        line = lines[pos]
        pos += 1
        if kind == "f":
            v = _numbers(line, float)
            if not v:
                raise ValueError(f"{path}: no value for {name}")
            vals[name] = v[0]
        elif kind == "i":
            v = _numbers(line, int)
            if not v:
                raise ValueError(f"{path}: no value for {name}")
            vals[name] = v[0]
        elif kind == "vf":
            vals[name] = _numbers(line, float)
        else:
            vals[name] = _numbers(line, int)
    return vals


@dataclass
class SysParams:
    diam: float = 0.0
    obs: float = 0.0
    tFrame: float = 0.0
    nW: int = 0
    nLgs: int = 0
    nTarget: int = 1
    nssp: int = 0
    gsAlt: List[float] = field(default_factory=list)
    gsType: List[int] = field(default_factory=list)
    alphaX_as: List[float] = field(default_factory=list)
    alphaY_as: List[float] = field(default_factory=list)
    XPup: List[float] = field(default_factory=list)
    YPup: List[float] = field(default_factory=list)
    thetaML: List[float] = field(default_factory=list)
    thetaCam: List[float] = field(default_factory=list)
    sensibilite: List[float] = field(default_factory=list)
    tracking: List[float] = field(default_factory=list)
    pasDphi: float = 1e-4
    ncpu: int = 1
    mrNGS: List[float] = field(default_factory=list)
    lgsFlux: float = 0.0
    ngsPixSize: float = 0.0
    lgsPixSize: float = 0.0
    lambdaNGS: float = 0.0
    lambdaLGS: float = 0.0
    bdw_m: float = 0.0
    throughNGS: float = 0.0
    throughLGS: float = 0.0
    throughAtm: float = 0.0
    RON: int = 0
    lgsCst: float = 0.0
    spotWidth: float = 0.0
    lgsAlt: float = 0.0
    lgsDepth: float = 0.0
    targetX_as: List[float] = field(default_factory=list)
    targetY_as: List[float] = field(default_factory=list)

    @property
    def nNgs(self) -> int:
        return self.nW - self.nLgs - 1

    def check(self):
        """Vector-size check of SysParams::vectorSize (sysParams.cpp:113-132): too short raises."""
        need = {"gsAlt": self.nW, "gsType": self.nW, "alphaX_as": self.nW, "alphaY_as": self.nW,
                "XPup": self.nW, "YPup": self.nW, "thetaML": self.nW, "thetaCam": self.nW,
                "sensibilite": self.nW, "tracking": 3, "mrNGS": self.nW - self.nLgs,
                "targetX_as": self.nTarget, "targetY_as": self.nTarget}
        for k, n in need.items():
            if len(getattr(self, k)) < n:
                raise ValueError(f"System Parameter vector size does not matches ({k})")


@dataclass
class AtmParams:
    nLayer: int = 0
    r0: float = 0.0
    cn2: List[float] = field(default_factory=list)
    h: List[float] = field(default_factory=list)
    L0: List[float] = field(default_factory=list)

    def check(self):
        for k in ("cn2", "h", "L0"):
            if len(getattr(self, k)) < self.nLayer:
                raise ValueError(f"Atmospheric Parameter vector size does not matches ({k})")


def read_sys(path: str) -> SysParams:
    s = SysParams(**_read(path, SYS_FIELDS))
    s.check()
    return s


def read_atm(path: str) -> AtmParams:
    a = AtmParams(**_read(path, ATM_FIELDS))
    a.check()
    return a
