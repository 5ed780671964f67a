"""CPU oracle for TeaLeaf.jl's linear-solve path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module; it is the checker,
never the thing measured or shipped.

The deck parser here restates ``Settings(infile)`` and ``readstate`` of the
reference (src/settings.jl:90-173) independently of the product's C++ parser,
so the two can be compared.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

LIT_D1, LIT_D2, LIT_D3, LIT_D4 = 1, 2, 4, 8
CG, CHEBY, PPCG, JACOBI = 0, 1, 2, 3
SOLVERS = {"cg": CG, "cheby": CHEBY, "ppcg": PPCG, "jacobi": JACOBI}
RECT, CIRC, POINT = 0, 1, 2
STOP_REFERENCE, STOP_RELATIVE = 0, 1
FIELDS = ["density0", "density", "energy0", "energy", "u", "u0", "p", "r", "mi", "w", "kx", "ky", "sd"]
FID = {n: i for i, n in enumerate(FIELDS)}
X_DENSITY, X_P, X_ENERGY0, X_ENERGY, X_U, X_SD = 1, 2, 4, 8, 16, 32
MAX_STATES = 128


class _State(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("density", "energy", "xmin", "ymin", "xmax", "ymax", "radius")] + [
        ("geometry", C.c_int)
    ]


class _Settings(C.Structure):
    _fields_ = [
        ("xcells", C.c_int), ("ycells", C.c_int), ("halodepth", C.c_int),
        ("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double),
        ("endstep", C.c_int), ("presteps", C.c_int), ("maxiters", C.c_int), ("coefficient", C.c_int),
        ("ppcginnersteps", C.c_int), ("summaryfrequency", C.c_int),
        ("errorswitch", C.c_int), ("checkresult", C.c_int),
        ("eps", C.c_double), ("dtinit", C.c_double), ("epslim", C.c_double),
        ("solver", C.c_int), ("nstates", C.c_int), ("states", _State * MAX_STATES),
        ("literal", C.c_uint), ("stop_rule", C.c_int), ("rel_tol", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("cg_iters", C.c_int), ("cheby_iters", C.c_int), ("est_iters", C.c_int),
        ("status", C.c_int), ("error", C.c_double), ("eigmin", C.c_double), ("eigmax", C.c_double),
        ("rel_resid", C.c_double),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(_Settings)]
        L.or_destroy.argtypes = [vp]
        L.or_dims.argtypes = [vp, ip, ip]
        L.or_set_maxiters.argtypes = [vp, C.c_int]
        L.or_field.restype = dp
        L.or_field.argtypes = [vp, C.c_int]
        L.or_history.restype = dp
        L.or_history.argtypes = [vp, C.c_int]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_initialise.argtypes = [vp]
        L.or_halo_update.argtypes = [vp, C.c_uint, C.c_int]
        L.or_cg_init.restype = C.c_double
        L.or_cg_init.argtypes = [vp, C.c_double, C.c_double, ip]
        L.or_cg_w.restype = C.c_double
        L.or_cg_w.argtypes = [vp]
        L.or_cg_ur.restype = C.c_double
        L.or_cg_ur.argtypes = [vp, C.c_double]
        L.or_cg_p.argtypes = [vp, C.c_double]
        L.or_smvp_field.argtypes = [vp, C.c_int, C.c_int]
        L.or_residual.argtypes = [vp]
        L.or_finalise.argtypes = [vp]
        L.or_field_summary.argtypes = [vp, dp]
        L.or_solve.argtypes = [vp, C.c_double, C.c_double, C.POINTER(Stats)]
        L.or_step.argtypes = [vp, C.POINTER(Stats)]
        L.or_run.argtypes = [vp, C.c_int, dp, C.POINTER(Stats)]
        L.or_tqli.argtypes = [dp, dp, C.c_int]
        L.or_eigenvalues.argtypes = [dp, dp, C.c_int, dp, dp]
        L.or_cheby_coef.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp, dp]
        L.or_calciter.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- deck parser


@dataclass
class State:
    density: float = 0.0
    energy: float = 0.0
    xmin: float = 0.0
    ymin: float = 0.0
    xmax: float = 0.0
    ymax: float = 0.0
    radius: float = 0.0
    geometry: int = RECT


@dataclass
class Deck:
    """Settings of src/settings.jl:39-78 (defaults as there)."""

    endstep: int = 2**63 - 1
    presteps: int = 30
    maxiters: int = 10_000
    coefficient: int = 1
    ppcginnersteps: int = 10
    summaryfrequency: int = 10
    halodepth: int = 2
    errorswitch: bool = False
    checkresult: bool = True
    eps: float = 1e-15
    dtinit: float = 0.1
    endtime: float = 10.0
    epslim: float = 1e-5
    solver: str = "cg"
    xcells: int = 10
    ycells: int = 10
    xmin: float = 0.0
    ymin: float = 0.0
    xmax: float = 100.0
    ymax: float = 100.0
    dx: float = 10.0
    dy: float = 10.0
    states: List[State] = field(default_factory=list)
    literal: int = 0
    stop_rule: int = STOP_REFERENCE
    rel_tol: float = 1e-10
    warnings: List[str] = field(default_factory=list)


_INT_KEYS = {"endstep", "presteps", "maxiters", "coefficient", "ppcginnersteps", "summaryfrequency",
             "halodepth", "xcells", "ycells"}
_FLOAT_KEYS = {"eps", "dtinit", "endtime", "epslim", "xmin", "ymin", "xmax", "ymax", "dx", "dy"}
_BOOL_KEYS = {"errorswitch", "checkresult"}


def _parse_bool(v: str) -> bool:
    v = v.strip()
    if v in ("true", "1"):
        return True
    if v in ("false", "0"):
        return False
    raise ValueError(v)


def parse_deck(text: str, literal: int = 0) -> Deck:
    """Settings(infile), src/settings.jl:90-135, with D22 (whitespace) corrected.

    State bounds are offset inwards by dx/100 and dy/100 (settings.jl:159-162)
    using the final dx, dy; with LIT_D3 the parse-time dx (default 10) is used
    as the reference does.
    """
    d = Deck(literal=literal)
    raw_states = []
    for line in text.splitlines():
        line = line.strip()
        if line.startswith("state"):
            raw_states.append((line, d.dx, d.dy))
            continue
        if line.startswith("use_"):
            name = line[4:].strip()
            if name in SOLVERS:
                d.solver = name
            else:
                d.warnings.append(f"Unknown setting {line}")
            continue
        if line.startswith("*") or "=" not in line:
            continue
        key, val = line.split("=", 1)
        key = key.replace("_", "").strip()
        if key == "initialtimestep":
            key = "dtinit"
        try:
            if key in _INT_KEYS:
                setattr(d, key, int(val.strip()))
            elif key in _FLOAT_KEYS:
                setattr(d, key, float(val.strip()))
            elif key in _BOOL_KEYS:
                setattr(d, key, _parse_bool(val))
            else:
                raise KeyError(key)
        except (KeyError, ValueError):
            d.warnings.append(f"Unknown setting {line}")
    d.dx = (d.xmax - d.xmin) / d.xcells
    d.dy = (d.ymax - d.ymin) / d.ycells
    for line, pdx, pdy in raw_states:
        d.states.append(_read_state(line, d, pdx if literal & LIT_D3 else d.dx, pdy if literal & LIT_D3 else d.dy))
    return d


def _read_state(line: str, d: Deck, dx: float, dy: float) -> State:
    """readstate, src/settings.jl:142-173."""
    parts = line.split()
    num = int(parts[1])
    s = State()
    for pair in parts[2:]:
        key, val = pair.split("=", 1)
        if key == "density":
            s.density = float(val)
        elif key == "energy":
            s.energy = float(val)
        if num == 1:
            continue
        if key == "xmin":
            s.xmin = float(val) + dx / 100
        elif key == "ymin":
            s.ymin = float(val) + dy / 100
        elif key == "xmax":
            s.xmax = float(val) - dx / 100
        elif key == "ymax":
            s.ymax = float(val) - dy / 100
        elif key == "radius":
            s.radius = float(val)
        elif key == "geometry":
            s.geometry = {"rectangle": RECT, "circular": CIRC, "point": POINT}[val]
    return s


def to_settings(d: Deck) -> _Settings:
    s = _Settings()
    for n in ("xcells", "ycells", "halodepth", "presteps", "maxiters", "coefficient", "ppcginnersteps",
              "summaryfrequency"):
        setattr(s, n, int(getattr(d, n)))
    s.endstep = int(min(d.endstep, 2**31 - 1))
    for n in ("xmin", "ymin", "xmax", "ymax", "dx", "dy", "eps", "dtinit", "epslim", "rel_tol"):
        setattr(s, n, float(getattr(d, n)))
    s.errorswitch = int(d.errorswitch)
    s.checkresult = int(d.checkresult)
    s.solver = SOLVERS[d.solver]
    if len(d.states) > MAX_STATES:
        raise ValueError("too many states")
    s.nstates = len(d.states)
    for i, st in enumerate(d.states):
        t = s.states[i]
        for n in ("density", "energy", "xmin", "ymin", "xmax", "ymax", "radius"):
            setattr(t, n, getattr(st, n))
        t.geometry = st.geometry
    s.literal = d.literal
    s.stop_rule = d.stop_rule
    return s


# ------------------------------------------------------------------- chunk


class Oracle:
    """A chunk of the CPU oracle, built from a parsed deck."""

    def __init__(self, deck: Deck, threads: Optional[int] = None):
        self.deck = deck
        self._s = to_settings(deck)
        L = lib()
        if threads:
            L.or_set_threads(int(threads))
        self._c = L.or_create(C.byref(self._s))
        if not self._c:
            raise MemoryError("or_create failed")
        X, Y = C.c_int(), C.c_int()
        L.or_dims(self._c, C.byref(X), C.byref(Y))
        self.X, self.Y = X.value, Y.value
        self.hd = deck.halodepth

    def __del__(self):
        c = getattr(self, "_c", None)
        if c:
            lib().or_destroy(c)
            self._c = None

    def field(self, name: str) -> np.ndarray:
        """(Y, X) numpy view: [j, k] = the reference's [k+1, j+1].  The view
        keeps this chunk alive."""
        p = lib().or_field(self._c, FID[name])
        buf = (C.c_double * (self.Y * self.X)).from_address(C.addressof(p.contents))
        buf._owner = self
        return np.frombuffer(buf, dtype=np.float64).reshape(self.Y, self.X)

    def interior(self, name: str) -> np.ndarray:
        h = self.hd
        return self.field(name)[h:self.Y - h, h:self.X - h]

    def history(self, which: int, n: int) -> np.ndarray:
        p = lib().or_history(self._c, which)
        return np.ctypeslib.as_array(p, shape=(n + 2,)).copy()

    def set_maxiters(self, n: int):
        if lib().or_set_maxiters(self._c, n):
            raise ValueError("maxiters above the allocated history")

    def initialise(self):
        rc = lib().or_initialise(self._c)
        if rc:
            raise RuntimeError(f"or_initialise failed ({rc})")

    def rxry(self):
        d = self.deck
        return d.dtinit / (d.dx * d.dx), d.dtinit / (d.dy * d.dy)

    def halo_update(self, mask: int, depth: int = 1):
        lib().or_halo_update(self._c, mask, depth)

    def cg_init(self):
        st = C.c_int()
        rx, ry = self.rxry()
        v = lib().or_cg_init(self._c, rx, ry, C.byref(st))
        if st.value:
            raise ValueError(f"Coefficient {self.deck.coefficient} is not valid")
        return v

    def cg_w(self):
        return lib().or_cg_w(self._c)

    def cg_ur(self, alpha):
        return lib().or_cg_ur(self._c, alpha)

    def cg_p(self, beta):
        lib().or_cg_p(self._c, beta)

    def smvp(self, src: str, dst: str):
        lib().or_smvp_field(self._c, FID[src], FID[dst])

    def residual(self):
        lib().or_residual(self._c)

    def finalise(self):
        lib().or_finalise(self._c)

    def summary(self) -> np.ndarray:
        out = np.zeros(4)
        lib().or_field_summary(self._c, _dp(out))
        return out

    def step(self) -> dict:
        st = Stats()
        rc = lib().or_step(self._c, C.byref(st))
        if rc:
            raise RuntimeError(f"or_step failed ({rc})")
        return st.as_dict()

    def run(self, nsteps: int):
        """nsteps timesteps; returns (summaries[nsteps, 4], list of stats dicts)."""
        summ = np.zeros((nsteps, 4))
        stats = (Stats * nsteps)()
        rc = lib().or_run(self._c, nsteps, _dp(summ), stats)
        if rc:
            raise RuntimeError(f"or_run failed ({rc})")
        return summ, [s.as_dict() for s in stats]


def tqli(d: np.ndarray, e: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    ee = np.zeros(len(d))
    ee[: len(e)] = e
    rc = lib().or_tqli(_dp(d), _dp(ee), len(d))
    if rc:
        raise RuntimeError(f"tqli failed ({rc})")
    return np.sort(d)


def eigenvalues(cga: np.ndarray, cgb: np.ndarray, n: int):
    """cga, cgb 1-based arrays (index 0 unused)."""
    a = np.ascontiguousarray(cga, dtype=np.float64)
    b = np.ascontiguousarray(cgb, dtype=np.float64)
    mn, mx = C.c_double(), C.c_double()
    rc = lib().or_eigenvalues(_dp(a), _dp(b), n, C.byref(mn), C.byref(mx))
    if rc:
        raise RuntimeError(f"eigenvalues failed ({rc})")
    return mn.value, mx.value


def cheby_coef(eigmin: float, eigmax: float, n: int):
    a = np.zeros(n + 2)
    b = np.zeros(n + 2)
    th = C.c_double()
    lib().or_cheby_coef(eigmin, eigmax, n, _dp(a), _dp(b), C.byref(th))
    return a, b, th.value


def calciter(eigmin, eigmax, error, bb) -> int:
    return lib().or_calciter(eigmin, eigmax, error, bb)


def run_deck(text: str, nsteps: Optional[int] = None, literal: int = 0, threads: Optional[int] = None,
             solver: Optional[str] = None, **overrides):
    d = parse_deck(text, literal=literal)
    if solver:
        d.solver = solver
    for k, v in overrides.items():
        setattr(d, k, v)
    o = Oracle(d, threads=threads)
    o.initialise()
    n = nsteps if nsteps is not None else d.endstep
    return o, o.run(n)
