"""tea.in decks for the benchmark configurations.

`reference_deck(n, steps)`: the three-state problem of the TeaLeaf decks
(background rho=100, e=1e-4; x in [0,1], y in [1,2] with rho=0.1, e=25;
x in [1,6], y in [1,2] with rho=0.1, e=0.1) on [0,10]^2.

`random_rect_deck(seed, n, ...)`: background rho=100, e=1e-4 plus `nrect`
axis-aligned rectangles drawn with SplitMix64 (a portable generator, so the
same seed gives the same deck everywhere): width, height ~ U[0.2, 3.0],
xmin ~ U[0, 10-w], ymin ~ U[0, 10-h], rho = 10^U[-2, 2], e ~ U[1e-4, 25].
Values are written with %.17g so the deck text is the interchange format.
"""
from __future__ import annotations

MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014)."""

    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))


def _header(n: int, steps: int, extra: str) -> str:
    return (f"x_cells={n}\ny_cells={n}\nxmin=0.0\nymin=0.0\nxmax=10.0\nymax=10.0\n"
            f"initial_timestep=0.004\nend_step={steps}\nsummary_frequency=1\nhalo_depth=2\ncoefficient=1\n"
            + extra)


def reference_deck(n: int = 512, steps: int = 10, solver: str = "cg", max_iters: int = 10000,
                   eps: float = 1e-15, presteps: int = 30, ppcg_inner_steps: int = 10, extra: str = "") -> str:
    return ("*tea\n"
            "state 1 density=100.0 energy=0.0001\n"
            "state 2 density=0.1 energy=25.0 geometry=rectangle xmin=0.0 xmax=1.0 ymin=1.0 ymax=2.0\n"
            "state 3 density=0.1 energy=0.1 geometry=rectangle xmin=1.0 xmax=6.0 ymin=1.0 ymax=2.0\n"
            + _header(n, steps, f"max_iters={max_iters}\neps={eps!r}\npresteps={presteps}\n"
                                f"ppcg_inner_steps={ppcg_inner_steps}\nuse_{solver}\n" + extra)
            + "*endtea\n")


def random_rect_deck(seed: int, n: int = 8192, steps: int = 20, nrect: int = 32, rel_tol: float = 1e-10,
                     max_iters: int = 100000, precision: str = "fp64", warm_start: str = "none",
                     extra: str = "") -> str:
    g = SplitMix64(seed)
    lines = ["*tea", "state 1 density=100.0 energy=0.0001"]
    for i in range(nrect):
        w = g.uniform(0.2, 3.0)
        h = g.uniform(0.2, 3.0)
        x0 = g.uniform(0.0, 10.0 - w)
        y0 = g.uniform(0.0, 10.0 - h)
        rho = 10.0 ** g.uniform(-2.0, 2.0)
        e = g.uniform(1e-4, 25.0)
        lines.append(f"state {i + 2} density={rho:.17g} energy={e:.17g} geometry=rectangle "
                     f"xmin={x0:.17g} xmax={x0 + w:.17g} ymin={y0:.17g} ymax={y0 + h:.17g}")
    body = _header(n, steps, f"max_iters={max_iters}\neps=1e-15\nuse_cg\nstop_rule=relative\n"
                             f"rel_tol={rel_tol!r}\nprecision={precision}\nwarm_start={warm_start}\n" + extra)
    return "\n".join(lines) + "\n" + body + "*endtea\n"
