"""Driver (claw2ez) cases shared by the driver tests and their golden
generator: run parameters written through ClawRunData, the initial data
and the solver of each case."""
import numpy as np

from classic_amd import problems as P
from classic_amd import rundata


def _common(cd, nx, ny, meqn, mwaves):
    cd.num_cells = [nx, ny]
    cd.num_eqn = meqn
    cd.num_waves = mwaves
    cd.num_aux = 0
    cd.capa_index = 0
    cd.t0 = 0.0
    cd.dt_max = 1e99
    cd.cfl_desired = 0.9
    cd.cfl_max = 1.0
    cd.steps_max = 5000
    cd.order = 2
    cd.transverse_waves = 2
    cd.dimensional_split = "unsplit"
    cd.use_fwaves = False
    cd.source_split = 0
    cd.num_ghost = 2


def acoustics_case(outstyle=1, nx=24, ny=20):
    """The radial acoustics pulse of examples/acoustics_2d_radial on
    [-1,1]^2 with rho = 1, bulk = 4, walls, MC."""
    rd = rundata.ClawRunData("classic", 2)
    prob = rd.new_UserData(name="probdata", fname="setprob.data")
    prob.add_param("rho", 1.0, "density of medium")
    prob.add_param("bulk", 4.0, "bulk modulus")
    cd = rd.clawdata
    cd.lower = [-1.0, -1.0]
    cd.upper = [1.0, 1.0]
    _common(cd, nx, ny, 3, 2)
    cd.output_style = outstyle
    if outstyle == 1:
        cd.num_output_times = 3
        cd.tfinal = 0.3
        cd.output_t0 = True
    elif outstyle == 2:
        cd.output_times = [0.05, 0.12, 0.3]
    else:
        cd.output_step_interval = 2
        cd.total_steps = 6
        cd.output_t0 = True
    cd.dt_initial = 0.01
    cd.dt_variable = True
    cd.limiter = ["mc", "mc"]
    cd.bc_lower = ["wall", "wall"]
    cd.bc_upper = ["wall", "wall"]
    q, _, _ = P.acoustics_pulse(nx, ny)
    return rd, q, "acoustics", (1.0, 4.0)


def euler_case(outstyle=3, nx=20, ny=16, seed=3):
    """Four-quadrant Euler on [0,1]^2, extrapolation, MC, one step per
    claw2 call (output_style 3) by default."""
    rd = rundata.ClawRunData("classic", 2)
    prob = rd.new_UserData(name="probdata", fname="setprob.data")
    prob.add_param("gamma", 1.4, "ratio of specific heats")
    cd = rd.clawdata
    cd.lower = [0.0, 0.0]
    cd.upper = [1.0, 1.0]
    _common(cd, nx, ny, 4, 3)
    cd.output_style = outstyle
    if outstyle == 1:
        cd.num_output_times = 2
        cd.tfinal = 0.08
        cd.output_t0 = True
    else:
        cd.output_step_interval = 2
        cd.total_steps = 6
        cd.output_t0 = True
    cd.dt_initial = 0.02  # too large at first: the first step is retaken
    cd.dt_variable = True
    cd.limiter = ["mc", "mc", "mc"]
    cd.bc_lower = ["extrap", "extrap"]
    cd.bc_upper = ["extrap", "extrap"]
    q, _, _ = P.euler_quadrants(nx, ny, seed=seed)
    return rd, q, "euler", (1.4, 1.0)


def acoustics1d_case(outstyle=1, mx=120):
    """examples/acoustics_1d_example1: Gaussian pressure pulse on [-1,1],
    rho = 1, bulk = 4, MC, extrapolation left and wall right."""
    rd = rundata.ClawRunData("classic", 1)
    prob = rd.new_UserData(name="probdata", fname="setprob.data")
    prob.add_param("rho", 1.0, "density of medium")
    prob.add_param("bulk", 4.0, "bulk modulus")
    cd = rd.clawdata
    cd.lower = [-1.0]
    cd.upper = [1.0]
    cd.num_cells = [mx]
    cd.num_eqn = 2
    cd.num_waves = 2
    cd.num_aux = 0
    cd.t0 = 0.0
    cd.output_style = outstyle
    if outstyle == 1:
        cd.num_output_times = 4
        cd.tfinal = 0.4
        cd.output_t0 = True
    else:
        cd.output_step_interval = 3
        cd.total_steps = 9
        cd.output_t0 = True
    cd.dt_initial = 0.05  # the first step is retaken
    cd.dt_max = 1e99
    cd.cfl_desired = 0.9
    cd.cfl_max = 1.0
    cd.steps_max = 5000
    cd.dt_variable = True
    cd.order = 2
    cd.limiter = ["mc", "mc"]
    cd.num_ghost = 2
    cd.bc_lower = ["extrap"]
    cd.bc_upper = ["wall"]
    q, _ = P.acoustics1d_gaussian(mx)
    return rd, q, "acoustics", (1.0, 4.0)


def acoustics3d_case(outstyle=1, n=(10, 8, 6)):
    """3D variable-coefficient acoustics in a two-layer medium (the
    heterogeneous example's setaux), Gaussian pulse, walls, method (2, 22),
    van Leer; aux written with frame 0 only."""
    rd = rundata.ClawRunData("classic", 3)
    cd = rd.clawdata
    cd.lower = [-1.0, -1.0, -1.0]
    cd.upper = [1.0, 1.0, 1.0]
    cd.num_cells = list(n)
    cd.num_eqn = 4
    cd.num_waves = 2
    cd.num_aux = 2
    cd.output_aux_components = "all"
    cd.output_aux_onlyonce = True
    cd.t0 = 0.0
    cd.output_style = outstyle
    if outstyle == 1:
        cd.num_output_times = 2
        cd.tfinal = 0.2
        cd.output_t0 = True
    else:
        cd.output_step_interval = 1
        cd.total_steps = 3
        cd.output_t0 = True
    cd.dt_initial = 0.01
    cd.dt_max = 1e99
    cd.cfl_desired = 0.9
    cd.cfl_max = 1.0
    cd.steps_max = 5000
    cd.dt_variable = True
    cd.order = 2
    cd.transverse_waves = 22
    cd.dimensional_split = "unsplit"
    cd.limiter = ["vanleer", "vanleer"]
    cd.num_ghost = 2
    cd.bc_lower = ["wall", "wall", "wall"]
    cd.bc_upper = ["wall", "wall", "wall"]
    mx, my, mz = n
    mbc = 2
    xs = [P.cell_centers(-1.0, 1.0, m, mbc)[0] for m in n]
    q = np.zeros((mz + 2 * mbc, my + 2 * mbc, mx + 2 * mbc, 4))
    r2 = ((xs[2][:, None, None] - 0.1) ** 2 + (xs[1][None, :, None] + 0.2) ** 2
          + (xs[0][None, None, :] - 0.15) ** 2)
    q[..., 0] = np.exp(-r2 / 0.2 ** 2)
    aux = P.layered_medium_aux(mx, my, mz, mbc=mbc, below=(1.0, 1.0), above=(6.0, 2.0))
    return rd, (q, aux), "vc_acoustics", ()


CASES = {
    "acoustics_style1": lambda: acoustics_case(1),
    "acoustics_style2": lambda: acoustics_case(2),
    "euler_style3": lambda: euler_case(3),
}

CASES1 = {
    "acoustics1d_style1": lambda: acoustics1d_case(1),
    "acoustics1d_style3": lambda: acoustics1d_case(3),
}

CASES3 = {
    "acoustics3d_style1": lambda: acoustics3d_case(1),
    "acoustics3d_style3": lambda: acoustics3d_case(3),
}

# cases whose reference outputs are committed under tests/golden/claw2ez_<name>
GOLDEN = ["acoustics_style1", "euler_style3"]
GOLDEN1 = ["acoustics1d_style1"]
GOLDEN3 = ["acoustics3d_style1"]


def qinit_from(q0):
    def qinit(*args):
        return np.array(q0, dtype=np.float64, copy=True)
    return qinit


def setaux_from(aux0):
    def setaux(*args):
        return np.array(aux0, dtype=np.float64, copy=True)
    return setaux
