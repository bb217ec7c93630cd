"""Reader/writer of the reference's positional input files.

sharpclaw.data is read by src/main.f90:61-122 after src/opendatafile.f:45-53
skips the '#' header and consumes the first non-'#' line (which must
therefore be blank).  Each later line holds one value (or, for mthlim, one
value per wave); text after the values is ignored by list-directed reads.
The same header convention holds for setprob.data (opendatafile again).
"""
from dataclasses import dataclass, field
from typing import List


@dataclass
class ClawData:
    ndim: int = 1
    nx: List[int] = field(default_factory=lambda: [100])
    nout: int = 1
    outstyle: int = 1
    tfinal: float = 1.0
    tout: List[float] = field(default_factory=list)   # outstyle 2
    nstepout: int = 1                                   # outstyle 3
    nstop: int = 1                                      # outstyle 3
    dt_initial: float = 1e-3
    dt_max: float = 1e99
    cfl_max: float = 1.0
    cfl_desired: float = 0.9
    max_steps: int = 50000
    dt_variable: int = 1
    time_integrator: int = 3
    verbosity: int = 0
    src_term: int = 0
    mcapa: int = 0
    maux: int = 0
    tfluct_solver: int = 0
    char_decomp: int = 0
    meqn: int = 1
    mwaves: int = 1
    lim_type: int = 2
    mthlim: List[int] = field(default_factory=lambda: [5])
    t0: float = 0.0
    xlower: List[float] = field(default_factory=lambda: [0.0])
    xupper: List[float] = field(default_factory=lambda: [1.0])
    mbc: int = 3
    mthbc: List[int] = field(default_factory=lambda: [1, 1])

    def write(self, path):
        r = repr
        L = ["# sharpclaw.data", "#", ""]
        L.append(f"{self.ndim}  =: ndim")
        for n in self.nx[: self.ndim]:
            L.append(f"{n}  =: nx")
        L.append("")
        L.append(f"{self.nout}  =: nout")
        L.append(f"{self.outstyle}  =: outstyle")
        if self.outstyle == 1:
            L.append(f"{r(float(self.tfinal))}  =: tfinal")
        elif self.outstyle == 2:
            L.append(" ".join(r(float(t)) for t in self.tout) + "  =: tout")
        else:
            L.append(f"{self.nstepout} {self.nstop}  =: nstepout nstop")
        L.append("")
        for name in ("dt_initial", "dt_max", "cfl_max", "cfl_desired"):
            L.append(f"{r(float(getattr(self, name)))}  =: {name}")
        L.append(f"{self.max_steps}  =: max_steps")
        L.append("")
        for name in ("dt_variable", "time_integrator", "verbosity", "src_term", "mcapa", "maux",
                     "tfluct_solver", "char_decomp"):
            L.append(f"{getattr(self, name)}  =: {name}")
        L.append("")
        L.append(f"{self.meqn}  =: meqn")
        L.append(f"{self.mwaves}  =: mwaves")
        L.append(f"{self.lim_type}  =: lim_type")
        L.append(" ".join(str(m) for m in self.mthlim) + "  =: mthlim")
        L.append("")
        L.append(f"{r(float(self.t0))}  =: t0")
        for d in range(self.ndim):
            L.append(f"{r(float(self.xlower[d]))}  =: xlower")
            L.append(f"{r(float(self.xupper[d]))}  =: xupper")
        L.append("")
        L.append(f"{self.mbc}  =: mbc")
        for m in self.mthbc[: 2 * self.ndim]:
            L.append(f"{m}  =: mthbc")
        L.append("0  =: restart")
        L.append("0  =: N_restart")
        with open(path, "w") as f:
            f.write("\n".join(L) + "\n")

    @classmethod
    def read(cls, path):
        """Positional read with main.f90:61-122's order (blank lines inside
        the value block are skipped like list-directed reads skip them)."""
        with open(path) as f:
            lines = f.read().splitlines()
        k = 0
        while k < len(lines) and lines[k].startswith("#"):
            k += 1
        k += 1  # opendatafile consumes the first non-'#' line
        vals = [ln.split("=:")[0].split() for ln in lines[k:]]
        vals = [v for v in vals if v]
        it = iter(vals)

        def one(conv=float):
            return conv(next(it)[0].replace("d", "e").replace("D", "e"))

        def many(n, conv=float):
            out = []
            while len(out) < n:
                out.extend(conv(x.replace("d", "e").replace("D", "e")) for x in next(it))
            return out[:n]

        d = cls()
        d.ndim = one(int)
        d.nx = [one(int) for _ in range(d.ndim)]
        d.nout = one(int)
        d.outstyle = one(int)
        if d.outstyle == 1:
            d.tfinal = one()
        elif d.outstyle == 2:
            d.tout = many(d.nout)
        else:
            d.nstepout, d.nstop = many(2, int)
        d.dt_initial = one()
        d.dt_max = one()
        d.cfl_max = one()
        d.cfl_desired = one()
        d.max_steps = one(int)
        d.dt_variable = one(int)
        d.time_integrator = one(int)
        d.verbosity = one(int)
        d.src_term = one(int)
        d.mcapa = one(int)
        d.maux = one(int)
        d.tfluct_solver = one(int)
        d.char_decomp = one(int)
        d.meqn = one(int)
        d.mwaves = one(int)
        d.lim_type = one(int)
        d.mthlim = many(d.mwaves, int)
        d.t0 = one()
        d.xlower, d.xupper = [], []
        for _ in range(d.ndim):
            d.xlower.append(one())
            d.xupper.append(one())
        d.mbc = one(int)
        d.mthbc = [one(int) for _ in range(2 * d.ndim)]
        return d


def write_values(path, values, title="data"):
    """A setprob.data-style file: '#' header, blank line, one value a line."""
    with open(path, "w") as f:
        f.write(f"# {title}\n#\n\n")
        for v in values:
            f.write((repr(float(v)) if isinstance(v, float) else str(v)) + "\n")
