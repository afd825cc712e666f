"""Host-side mirror of the reference's module global (src/module_global.f90).

`TCSConfig` holds the namelist inputs of PARA_INPUT.dat (module_global.f90:
83-105) and derives the constants the reference hard-codes or computes in
global_inputpara (:107-165).  The property polynomials, hard-coded in the
reference, are a field here so configurations can choose them.
"""
import dataclasses
import math
import re

import numpy as np


def _f32(x):
    """Value of a Fortran default-REAL literal stored into a double."""
    return float(np.float32(x))


@dataclasses.dataclass
class Properties:
    """Coefficients of the property polynomials in DeltaT*T:
    rho (a1x), Cp (b1x), k (c1x), nu (d1x) -- src/module_global.f90:112-132."""
    a: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    b: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    c: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    d: tuple = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0)

    @classmethod
    def constant(cls):
        """The active set of the reference (:117-120): all properties constant."""
        return cls()

    @classmethod
    def water(cls):
        """The reference's commented water set (:113-116).  Literals without a
        d exponent are Fortran default REAL (single precision) there, and are
        kept as such."""
        return cls(
            a=(0.9922e+3, -3.736e-4, -3.98e-6, 0.0, 0.0, 0.0),
            b=(4.1690e+3, 0.084e-4, 4.60e-6, 0.0, 0.0, 0.0),
            c=(_f32(0.6297), 21.99e-4, -17.8e-6, 0.0, 0.0, 0.0),
            d=(0.6690e-6, -175.9e-4, 295.8e-6, -460.e-8, 0.0, 0.0),
        )

    @classmethod
    def glycerol(cls):
        """The reference's commented glycerol set (:128-132)."""
        return cls(
            a=(1.2477e+3, -4.789e-4, -0.3795e-6, 0.0, 0.0, 0.0),
            b=(2.5108e+3, 22.511e-4, 0.0, 0.0, 0.0, 0.0),
            c=(2.9351e-3, 3.863e-4, 0.0, 0.0, 0.0, 0.0),
            d=(238.71e-6, -702.83e-4, 2393.1e-6, -6923.0e-8, 33131.3e-10, -71517.5e-12),
        )


@dataclasses.dataclass
class TCSConfig:
    n1m: int = 64
    n2m: int = 64
    n3m: int = 64
    np1: int = 1
    np2: int = 1
    np3: int = 1
    pbc1: bool = True
    pbc2: bool = False
    pbc3: bool = True
    uniform1: int = 1
    uniform2: int = 1
    uniform3: int = 1
    gamma1: float = 3.0
    gamma2: float = 2.6
    gamma3: float = 3.0
    Aspect1: float = math.pi
    H: float = 2.0
    Aspect3: float = math.pi
    Ra: float = 1.0e6
    Pr: float = 0.71
    DeltaT: float = 10.0
    MaxCFL: float = 0.5
    dtStart: float = 1.0e-3
    tStart: float = 0.0
    Timestepmax: int = 100
    props: Properties = dataclasses.field(default_factory=Properties.constant)
    # wall values (module_global.f90:107-110)
    Uup: float = 0.0
    Ubt: float = 0.0
    Vup: float = 0.0
    Vbt: float = 0.0
    Wup: float = 0.0
    Wbt: float = 0.0
    Tup: float = -0.5
    Tbt: float = 0.5

    # ---- derived constants (global_inputpara, :122-165) -----------------
    @property
    def a12pera11(self):
        a11, a12 = self.props.a[1], self.props.a[2]
        return 0.0 if a12 < 1.0e-14 else a12 / a11

    @property
    def Cmu(self):
        return math.sqrt(self.Pr / self.Ra)

    @property
    def Cmt(self):
        return 1.0

    @property
    def Ct(self):
        return 1.0 / math.sqrt(self.Ra * self.Pr)

    @property
    def Cmp(self):
        return 1.0

    @property
    def L1(self):
        return self.H * self.Aspect1

    @property
    def L2(self):
        return self.H

    @property
    def L3(self):
        return self.H * self.Aspect3

    @property
    def dtMax(self):
        t = [self.L1 / float(self.n1m), self.L2 / float(self.n2m), self.L3 / float(self.n3m)]
        return min(t) * 100

    def constants(self):
        """Name -> value of every physical constant the kernels use."""
        a, b, c, d = self.props.a, self.props.b, self.props.c, self.props.d
        out = {"Ct": self.Ct, "Cmu": self.Cmu, "Cmt": self.Cmt, "Cmp": self.Cmp, "DeltaT": self.DeltaT}
        for name, coef in zip("abcd", (a, b, c, d)):
            for q in range(6):
                out[f"{name}1{q}"] = float(coef[q])
        out.update(KPP0=float(c[0]), Rho0=float(a[0]), Cp0=float(b[0]), Mu0=float(a[0]) * float(d[0]),
                   a12pera11=self.a12pera11, L1=self.L1, L2=self.L2, L3=self.L3,
                   MaxCFL=self.MaxCFL, dtMax=self.dtMax)
        return out

    # ---- namelist input ---------------------------------------------------
    @classmethod
    def from_namelist(cls, path, props=None):
        """Read a PARA_INPUT.dat (module_global.f90:95-105)."""
        groups = parse_namelist(open(path).read())
        kw = {}
        flat = {}
        for g in groups.values():
            flat.update(g)
        keymap = {f.name.lower(): f.name for f in dataclasses.fields(cls)}
        for k, v in flat.items():
            if k in keymap:
                kw[keymap[k]] = v
        if props is not None:
            kw["props"] = props
        for key in ("pbc1", "pbc2", "pbc3"):
            if key in kw:
                kw[key] = bool(kw[key])
        return cls(**kw)


def _parse_value(tok):
    t = tok.strip().rstrip(",")
    tl = t.lower()
    if tl in (".true.", "t", ".t."):
        return True
    if tl in (".false.", "f", ".f."):
        return False
    if (t.startswith("'") and t.endswith("'")) or (t.startswith('"') and t.endswith('"')):
        return t[1:-1]
    try:
        return int(t)
    except ValueError:
        pass
    return float(tl.replace("d", "e"))


def parse_namelist(text):
    """Minimal Fortran namelist reader: {group: {key: value}} (keys lowered)."""
    groups = {}
    for m in re.finditer(r"&(\w+)(.*?)\n\s*/", text, flags=re.S):
        body = {}
        for line in m.group(2).splitlines():
            line = line.split("!")[0].strip()
            if "=" not in line:
                continue
            k, v = line.split("=", 1)
            body[k.strip().lower()] = _parse_value(v)
        groups[m.group(1).lower()] = body
    return groups
