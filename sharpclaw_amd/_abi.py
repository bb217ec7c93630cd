"""ctypes mirror of include/sharpclaw_amd.h (struct layouts and constants).

The parameter block carries what module ClawParams (src/ClawParams.f90:4-14)
and the physics common blocks (/comrp/, /param/, /cparam/) carry in the
reference.
"""
import ctypes as C

SC_OK = 0
SC_ERR_ARG = 1
SC_ERR_UNSUPPORTED = 2
SC_ERR_HIP = 3
SC_ERR_NOMEM = 4
SC_ERR_STATE = 5

SC_PHYS_ADVECTION = 1
SC_PHYS_ACOUSTICS = 2
SC_PHYS_EULER = 3
SC_PHYS_SHALLOW_WATER = 4

SC_FP64 = 0
SC_FP32 = 1

PHYSICS = {
    "advection": SC_PHYS_ADVECTION,
    "acoustics": SC_PHYS_ACOUSTICS,
    "euler": SC_PHYS_EULER,
    "shallow_water": SC_PHYS_SHALLOW_WATER,
}


class ScParams(C.Structure):
    _fields_ = [
        ("ndim", C.c_int),
        ("nx", C.c_int * 2),
        ("meqn", C.c_int),
        ("mwaves", C.c_int),
        ("mbc", C.c_int),
        ("maux", C.c_int),
        ("mcapa", C.c_int),
        ("lim_type", C.c_int),
        ("char_decomp", C.c_int),
        ("tfluct_solver", C.c_int),
        ("time_integrator", C.c_int),
        ("src_term", C.c_int),
        ("mthbc", C.c_int * 4),
        ("xlower", C.c_double * 2),
        ("dx", C.c_double * 2),
        ("physics", C.c_int),
        ("efix", C.c_int),
        ("gamma", C.c_double),
        ("grav", C.c_double),
        ("u", C.c_double),
        ("rho", C.c_double),
        ("bulk", C.c_double),
        ("precision", C.c_int),
        ("device", C.c_int),
        ("mthlim", C.c_int * 4),
        ("tperiod", C.c_double),
    ]


class ScRunCtl(C.Structure):
    _fields_ = [
        ("dt_max", C.c_double),
        ("cfl_max", C.c_double),
        ("cfl_desired", C.c_double),
        ("max_steps", C.c_int),
        ("dt_variable", C.c_int),
    ]


class ScRunStats(C.Structure):
    _fields_ = [
        ("t", C.c_double),
        ("cfl_maxused", C.c_double),
        ("cfl_last", C.c_double),
        ("dt_minused", C.c_double),
        ("dt_maxused", C.c_double),
        ("dt_last", C.c_double),
        ("steps_taken", C.c_int),
        ("info", C.c_int),
        ("rejected", C.c_int),
        ("stage_evals", C.c_longlong),
        ("gpu_ms", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ScBatchResult(C.Structure):
    _fields_ = [
        ("t", C.c_double),
        ("cfl_maxused", C.c_double),
        ("cfl_last", C.c_double),
        ("dt_minused", C.c_double),
        ("dt_maxused", C.c_double),
        ("dt_last", C.c_double),
        ("steps_taken", C.c_int),
        ("rejected", C.c_int),
        ("info", C.c_int),
        ("buffer", C.c_int),
        ("stage_evals", C.c_longlong),
        ("t_begin", C.c_longlong),
        ("t_end", C.c_longlong),
        ("done_ms", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ScRequestStats(C.Structure):
    _fields_ = [
        ("run", ScRunStats),
        ("submit_ms", C.c_double),
        ("admit_ms", C.c_double),
        ("done_ms", C.c_double),
        ("footprint", C.c_longlong),
        ("rounds", C.c_int),
    ]

    def as_dict(self):
        d = self.run.as_dict()
        d.update({f: getattr(self, f) for f, _ in self._fields_[1:]})
        return d


class ScServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_longlong),
        ("completed", C.c_longlong),
        ("rounds", C.c_longlong),
        ("peak_bytes", C.c_longlong),
        ("capacity_bytes", C.c_longlong),
        ("used_bytes", C.c_longlong),
        ("active", C.c_int),
        ("queued", C.c_int),
        ("peak_active", C.c_int),
        ("clock_ms", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# meqn / mwaves of each device physics by dimension
SHAPES = {
    (SC_PHYS_ADVECTION, 1): (1, 1),
    (SC_PHYS_ACOUSTICS, 1): (2, 2),
    (SC_PHYS_EULER, 1): (3, 3),
    (SC_PHYS_EULER, 2): (4, 3),
    (SC_PHYS_ADVECTION, 2): (1, 1),
    (SC_PHYS_SHALLOW_WATER, 1): (2, 2),
    (SC_PHYS_SHALLOW_WATER, 2): (3, 3),
}


def make_params(physics, nx, xlower=(0.0, 0.0), xupper=(1.0, 1.0), *, char_decomp=2,
                time_integrator=3, mthbc=(1, 1, 1, 1), gamma=1.4, grav=9.81, u=1.0, rho=1.0,
                bulk=4.0, precision=SC_FP64, efix=-1, device=0, tfluct_solver=0, lim_type=2,
                mthlim=(5, 5, 5, 5), maux=0, mcapa=0, src_term=0, tperiod=0.0):
    """Build an ScParams the way main.f90:61-122 fills ClawParams.

    dx = (xupper - xlower) / nx as main.f90:116 (float(nx) is exact here).
    """
    if isinstance(physics, str):
        physics = PHYSICS[physics]
    nx = tuple(nx) if isinstance(nx, (tuple, list)) else (nx,)
    ndim = len(nx)
    meqn, mwaves = SHAPES[(physics, ndim)]
    p = ScParams()
    p.ndim = ndim
    p.nx[0] = nx[0]
    p.nx[1] = nx[1] if ndim == 2 else 1
    p.meqn = meqn
    p.mwaves = mwaves
    p.mbc = 3
    p.maux = maux
    p.mcapa = mcapa
    p.lim_type = lim_type
    p.char_decomp = char_decomp
    p.tfluct_solver = tfluct_solver
    p.time_integrator = time_integrator
    p.src_term = src_term
    for k in range(4):
        p.mthbc[k] = mthbc[k] if k < len(mthbc) else 1
    for d in range(2):
        if d < ndim:
            p.xlower[d] = float(xlower[d])
            p.dx[d] = (float(xupper[d]) - float(xlower[d])) / float(nx[d])
        else:
            p.xlower[d] = 0.0
            p.dx[d] = 1.0
    p.physics = physics
    p.efix = efix
    p.gamma = gamma
    p.grav = grav
    p.u = u
    p.rho = rho
    p.bulk = bulk
    p.precision = precision
    p.device = device
    for k in range(4):
        p.mthlim[k] = mthlim[k] if k < len(mthlim) else mthlim[-1]
    p.tperiod = tperiod
    return p


def grid_shape(p):
    """numpy shape (Fortran order) of q(1-mbc:nx+mbc[,1-mbc:ny+mbc], meqn)."""
    if p.ndim == 1:
        return (p.nx[0] + 2 * p.mbc, p.meqn)
    return (p.nx[0] + 2 * p.mbc, p.nx[1] + 2 * p.mbc, p.meqn)


def interior(p, q):
    """View of the interior cells of a grid array."""
    b = p.mbc
    if p.ndim == 1:
        return q[b:b + p.nx[0], :]
    return q[b:b + p.nx[0], b:b + p.nx[1], :]
