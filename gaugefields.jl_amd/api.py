"""Host-side mirror of Gaugefields.jl's user API for the SU(3) hot path, on libgfhip.

Names, argument meaning and error behaviour follow the reference (Julia ``!`` functions
are the same names without ``!``; they mutate their first argument in place):

  gauge_configuration        API.jl:178-251
  measure_plaquette          API.jl:395-404       calculate_Plaquette  AbstractGaugefields.jl:2684-2699
  gauge_momenta              API.jl:330-331       gaussian_momenta(!)  API.jl:341-392
  GaugeAction / push         GaugeActions.jl:25-62, 238-246
  evaluate_GaugeAction       GaugeActions.jl:132-142
  calc_dSdUmu                GaugeActions.jl:77-123
  md_action_workspace / md_potential / md_force   molecular_dynamics.jl:13-43, 243-267
  md_driver / md_step_size / md_hamiltonian       molecular_dynamics.jl:425-505
  update_gaugefields / update_momenta             molecular_dynamics.jl:513-551
  QPQ / PQP / OmelyanPQP / md_step / md_trajectory molecular_dynamics.jl:180-200, 594-730;
                                                   docs/src/howtoimplement.md:94-112
  gradient_flow / flow       API.jl:420-437, gradientflow.jl:118-238
  stout_smearing / smear / back_prop   API.jl:452-480, Abstractsmearing.jl:237-411
  topological_charge(_density)  AbstractGaugefields.jl:1430-1471
  calculate_energy_density   samples/measurements/energydensity.jl:32-76 (clover)
  Heatbath / Heatbath_update / heatbath / overrelaxation   heatbath/heatbathmodule.jl:50-200, 539-602, 2119-2163
  ILDG / load_gaugefield / save_binarydata     ildg_format.jl:686-695, 797-833, 902-969
  save_configuration / load_configuration_     API.jl:711-727, 781-826 (format=:ildg)

Only what the hot path needs exists here: SU(3), four dimensions, periodic boundaries,
the Wilson plaquette action (loop set plaquette ∪ plaquette^dag).  Everything runs
through the HIP library; there is no CPU fallback.
"""
import math

import numpy as np

from . import generators
from .lattice import OMELYAN_LAMBDA, Lattice

__all__ = [
    "ArgumentError",
    "HIPBackend",
    "HIPGaugeConfiguration",
    "HIPMomenta",
    "gauge_configuration",
    "measure_plaquette",
    "calculate_Plaquette",
    "gauge_momenta",
    "gaussian_momenta",
    "gaussian_momenta_",
    "GaugeAction",
    "make_loops_fromname",
    "evaluate_GaugeAction",
    "calc_dSdUmu",
    "md_action_workspace",
    "md_potential",
    "md_force",
    "QPQ",
    "PQP",
    "OmelyanPQP",
    "MDDriver",
    "md_driver",
    "md_step_size",
    "md_hamiltonian",
    "update_gaugefields",
    "update_momenta",
    "md_step",
    "md_trajectory",
    "Gradientflow",
    "gradient_flow",
    "flow",
    "energy_density",
    "StoutSmearing",
    "stout_smearing",
    "smear",
    "back_prop",
    "SmearedGaugeAction",
    "topological_charge",
    "topological_charge_density",
    "calculate_energy_density",
    "Heatbath",
    "Heatbath_update",
    "heatbath",
    "overrelaxation",
    "ILDG",
    "load_gaugefield",
    "save_binarydata",
    "save_configuration",
    "load_configuration_",
]


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class HIPBackend:
    """Backend selector for gauge_configuration (the reference's AbstractGaugeBackend seam)."""


# ----------------------------------------------------------------------------- fields


class HIPGaugeConfiguration:
    """The four link fields U[1..4] of one configuration, resident on the GPU.

    ``len(U) == 4``; ``U.NC == 3``; ``U.NV`` sites.  The momenta of the configuration share its
    device context (``gauge_momenta(U)``)."""

    NC = 3
    Dim = 4

    def __init__(self, lattice: Lattice):
        self.lattice = lattice
        self.NV = lattice.V
        self.dims = lattice.dims
        self.smearing = None

    def __len__(self):
        return 4

    @property
    def eltype(self):
        # the links as held: FP64 also in the FP32-force form, whose forces alone are FP32
        return np.complex64 if self.lattice.precision == "f32" else np.complex128

    def to_host(self):
        """links as (4, V, 9) complex128, reference layout"""
        return self.lattice.download_links()

    def from_host(self, links):
        self.lattice.upload_links(links)

    def copy_to_backup(self):
        """substitute_U!(Uold, U)"""
        self.lattice.save_links()

    def restore_from_backup(self):
        """substitute_U!(U, Uold)"""
        self.lattice.restore_links()


class HIPMomenta:
    """Conjugate momenta (8 su(3) coefficients per link) living in the configuration's context."""

    def __init__(self, U: HIPGaugeConfiguration):
        self.U = U
        self.lattice = U.lattice

    def __len__(self):
        return 4

    def to_host(self):
        return self.lattice.download_momenta()

    def from_host(self, P):
        self.lattice.upload_momenta(P)

    def __mul__(self, other):
        """p*p (TA_gaugefields_4D_MPILattice.jl:147-155)"""
        if other is not self:
            raise ArgumentError("p*q is only provided for q === p on the HIP backend")
        return self.lattice.momenta_dot()


def _check_same(U, P):
    if not isinstance(P, HIPMomenta) or P.lattice is not U.lattice:
        raise ArgumentError("U and P must belong to the same HIP configuration")


def gauge_configuration(lattice, *, backend=None, colors=3, halo=1, start="cold", seed=None,
                        eltype=np.complex128, boundary="periodic", verbose=0, device=0, force_eltype=None):
    """Create a configuration on the GPU (gauge_configuration, API.jl:178-251).

    start="cold": unit links.  start="hot": the hot fill of the legacy backend (entries
    (u - 1/2) + i (u' - 1/2), rows Gram-Schmidt normalized to SU(3)) generated on the device from
    the Philox4x32-10 stream of ``seed``; ``seed=None`` draws a fresh seed.

    force_eltype (HIP backend only, no counterpart in the reference): np.complex64 with
    eltype=np.complex128 keeps the links and momenta in FP64 and evaluates every force in FP32
    (the "f32force" form)."""
    if backend is not None and not isinstance(backend, HIPBackend):
        raise ArgumentError("only HIPBackend() is available here")
    dims = tuple(int(d) for d in lattice)
    if len(dims) != 4:
        raise ArgumentError(f"the HIP backend is four dimensional; got {len(dims)} extents")
    if any(d <= 0 for d in dims):
        raise ArgumentError(f"lattice extents must be positive; got {dims}")
    if colors != 3:
        raise ArgumentError(f"the HIP backend implements SU(3) only; got colors={colors}")
    if start not in ("cold", "hot"):
        raise ArgumentError(f"start must be :cold or :hot; got {start}")
    if int(halo) < 0:
        raise ArgumentError(f"halo must be nonnegative; got {halo}")
    if boundary != "periodic":
        raise ArgumentError("the HIP backend supports periodic boundaries only")
    if eltype not in (np.complex128, np.complex64):
        raise ArgumentError("eltype must be ComplexF64 (np.complex128) or ComplexF32 (np.complex64)")
    if force_eltype not in (None, eltype, np.complex64):
        raise ArgumentError("force_eltype must be the links' eltype or ComplexF32 (np.complex64)")
    precision = "f64" if eltype == np.complex128 else "f32"
    if precision == "f64" and force_eltype == np.complex64:
        precision = "f32force"
    L = Lattice(dims, precision=precision, device=device)
    U = HIPGaugeConfiguration(L)
    if start == "hot":
        if seed is None:
            seed = int(np.random.SeedSequence().entropy % (2 ** 63))
        L.hot_start(int(seed) % (2 ** 64))
    return U


def calculate_Plaquette(U):
    """sum_{x, mu<nu} Re tr P_munu (the unnormalized historical value)"""
    return U.lattice.plaquette_sum()


def measure_plaquette(U, *, normalize=True):
    v = calculate_Plaquette(U)
    return v / (6 * U.NV * 3) if normalize else v


def gauge_momenta(U):
    """zero-valued momenta compatible with U"""
    P = HIPMomenta(U)
    U.lattice.clear_momenta()
    return P


def gaussian_momenta_(momenta, *, sigma=1.0, seed=None, sweep=0):
    """Fill momenta with N(0, sigma^2) coefficients (gaussian_momenta!, API.jl:341-368), generated
    on the device from the Philox4x32-10 per-link stream (seed, sweep) -- not the reference's
    LatticeMatrices stream."""
    if sweep < 0:
        raise ArgumentError(f"sweep must be nonnegative; got {sweep}")
    if not sigma >= 0:
        raise ArgumentError(f"sigma must be nonnegative; got {sigma}")
    if seed is None:
        seed = int(np.random.SeedSequence().entropy % (2 ** 63))
    momenta.lattice.gaussian_momenta(int(seed) % (2 ** 64), int(sweep), sigma=float(sigma))
    return momenta


def gaussian_momenta(U, *, sigma=1.0, seed=None, sweep=0):
    return gaussian_momenta_(HIPMomenta(U), sigma=sigma, seed=seed, sweep=sweep)


# ----------------------------------------------------------------------------- actions


_PLAQ = "plaquette"
_PLAQ_DAG = "plaquette'"


def make_loops_fromname(name, Dim=4):
    """Loop-set token for make_loops_fromname("plaquette"); `loops + dagger(loops)` closes it."""
    if name != "plaquette":
        raise ArgumentError(f"only the plaquette loop set is implemented on the HIP backend; got {name!r}")
    if Dim != 4:
        raise ArgumentError("Dim must be 4")
    return [_PLAQ]


def dagger(loops):
    """loops' (the Hermitian-conjugate paths)"""
    return [{_PLAQ: _PLAQ_DAG, _PLAQ_DAG: _PLAQ}[x] for x in loops]


class GaugeAction:
    """Wilson-loop action restricted to coefficient * (plaquette ∪ plaquette^dag) terms."""

    def __init__(self, U, hascovnet=False):
        self.U = U
        self.terms = []  # list of coefficients

    def push(self, coefficient, closedloops):
        loops = sorted(closedloops)
        if loops != sorted([_PLAQ, _PLAQ_DAG]):
            raise ArgumentError(
                "the HIP backend evaluates exactly the loop set plaquette ∪ plaquette' "
                f"(got {closedloops}); other Wilson loops are not on this path")
        self.terms.append(complex(coefficient))
        return self

    @property
    def coefficient(self):
        c = sum(self.terms)
        if c.imag != 0:
            raise ArgumentError("the HIP backend needs a real action coefficient")
        return c.real


def evaluate_GaugeAction(action, U):
    """traced, lattice-summed loop expression: 2 * coeff * sum Re tr P"""
    return U.lattice.action(action.coefficient)


def evaluate_GaugeAction_untraced(action, U):
    """the untraced loop sum of the action at every site (evaluate_GaugeAction_untraced,
    GaugeActions.jl:144-213): coeff sum_{mu<nu} (P_munu + P_munu^dag), (V, 3, 3) complex [i, j];
    its trace summed over sites is evaluate_GaugeAction"""
    return U.lattice.action_untraced(action.coefficient)


def calc_dSdUmu(action, mu, U):
    """dS/dU_mu (3x3 per site) for mu in 1..4, as (V, 3, 3) complex128 [i, j]"""
    if mu not in (1, 2, 3, 4):
        raise ArgumentError(f"mu must be in 1:4; got {mu}")
    D = U.lattice.dsdu(action.coefficient)
    return generators.links_to_matrices(D)[mu - 1]


class _Workspace:
    pass


class MDActionSet:
    """Independently scheduled MD action providers by name (MDActionSet, molecular_dynamics.jl:45-68):
    the potential and force are the sums of the members'.  Members are GaugeAction or
    SmearedGaugeAction providers; the names select them in MDForceGroup / SextonWeingarten."""

    def __init__(self, terms=None, **kw):
        terms = dict(terms or {}, **kw)
        if not terms:
            raise ArgumentError("an MDActionSet must contain at least one action provider")
        self.terms = terms

    def keys(self):
        return tuple(self.terms)


class MDForceGroup:
    """Named members of an MDActionSet for a momentum update (MDForceGroup, molecular_dynamics.jl:70-93)."""

    def __init__(self, *names):
        if len(names) == 1 and isinstance(names[0], (tuple, list)):
            names = tuple(names[0])
        if not names:
            raise ArgumentError("an MDForceGroup must contain at least one action name")
        if len(set(names)) != len(names):
            raise ArgumentError(f"an MDForceGroup must not contain duplicate action names: {tuple(names)}")
        self.names = tuple(names)


def _md_force_group(x):
    if isinstance(x, MDForceGroup):
        return x
    if isinstance(x, str):
        return MDForceGroup(x)
    if isinstance(x, (tuple, list)):
        return MDForceGroup(*x)
    raise ArgumentError(f"a force group is an action name, a tuple of names or an MDForceGroup; got {x!r}")


def _providers(action, group=None):
    """the (provider) members a force evaluation sums: the action itself, or the set's members
    (all, or those named by group, in the group's order)"""
    if isinstance(action, MDActionSet):
        names = group.names if group is not None else action.keys()
        return [action.terms[n] for n in names]
    if group is not None:
        raise ArgumentError(f"a force group selects members of an MDActionSet; got {type(action).__name__}")
    return [action]


def _configure_member(member, U):
    """the device's stout setting for one provider; returns its plaquette coefficient"""
    if isinstance(member, SmearedGaugeAction):
        member.configure(U)
    elif isinstance(member, GaugeAction):
        U.lattice.stout_config(0, 0.0)
    else:
        raise ArgumentError(f"md_force! is not implemented for {type(member).__name__}")
    return member.coefficient


def md_action_workspace(action, U):
    for m in _providers(action):
        if not isinstance(m, (GaugeAction, SmearedGaugeAction)):
            raise ArgumentError(f"md_action_workspace is not implemented for {type(m).__name__}")
    return _Workspace()


def _member_potential(member, U):
    if isinstance(member, SmearedGaugeAction):
        return member.potential(U)
    if not isinstance(member, GaugeAction):
        raise ArgumentError(f"md_potential is not implemented for {type(member).__name__}")
    return -evaluate_GaugeAction(member, U) / U.NC


def md_potential(action, U, workspace=None):
    """-Re S / NC of the action (molecular_dynamics.jl:247-249); the sum over the members of an
    MDActionSet (:111-126)"""
    return sum(_member_potential(m, U) for m in _providers(action))


def md_force(force, action, U, workspace=None, group=None):
    """Write dp/dtau into the configuration's force field; returns it as (4, V, 8).  For an
    MDActionSet the sum of the members' forces, or of those the MDForceGroup `group` names
    (molecular_dynamics.jl:201-236): the first member's force written, every further one added."""
    members = _providers(action, group)
    for k, m in enumerate(members):
        coeff = _configure_member(m, U)
        if k == 0:
            U.lattice.force(coeff)
        else:
            U.lattice.force_add(coeff)
    F = U.lattice.download_force()
    if force is not None:
        force[...] = F
    return F


# ----------------------------------------------------------------------------- MD


class QPQ:
    """Q(1/2) P(1) Q(1/2) (molecular_dynamics.jl:190-198)"""

    code = "QPQ"


class PQP:
    """P(1/2) Q(1) P(1/2) (molecular_dynamics.jl:180-188)"""

    code = "PQP"


class OmelyanPQP:
    """P(l) Q(1/2) P(1-2l) Q(1/2) P(l) (docs/src/howtoimplement.md:94-112)"""

    code = "2MN"

    def __init__(self, lam=OMELYAN_LAMBDA):
        self.lam = float(lam)


class SextonWeingarten:
    """Second-order Sexton-Weingarten integrator for two force time scales (molecular_dynamics.jl:
    298-360): slow and fast are action names, tuples of names or MDForceGroups of an MDActionSet.
    QPQ ordering: a fast evolution of half the step (n_fast QPQ substeps) on each side of one full
    slow kick; PQP ordering: half slow kicks around one full fast evolution."""

    def __init__(self, *, slow, fast, n_fast, ordering=None):
        ordering = QPQ() if ordering is None else ordering
        if int(n_fast) <= 0:
            raise ArgumentError(f"n_fast must be positive; got {n_fast}")
        if not isinstance(ordering, (PQP, QPQ)):
            raise ArgumentError(f"SextonWeingarten ordering must be PQP() or QPQ(); got {type(ordering).__name__}")
        self.slow = _md_force_group(slow)
        self.fast = _md_force_group(fast)
        self.n_fast = int(n_fast)
        self.ordering = ordering


def _validate_md_integrator(integrator, action):
    """molecular_dynamics.jl:362-407"""
    if not isinstance(integrator, SextonWeingarten):
        return
    if not isinstance(action, MDActionSet):
        raise ArgumentError(f"SextonWeingarten requires an MDActionSet; got {type(action).__name__}")
    available = action.keys()
    slow, fast = integrator.slow.names, integrator.fast.names
    overlap = tuple(n for n in slow if n in fast)
    if overlap:
        raise ArgumentError(f"slow and fast force groups overlap: {overlap}")
    unknown = tuple(n for n in slow + fast if n not in available)
    if unknown:
        raise ArgumentError(f"SextonWeingarten selects unknown actions {unknown}; available names are {available}")
    omitted = tuple(n for n in available if n not in slow + fast)
    if omitted:
        raise ArgumentError(f"SextonWeingarten does not schedule actions {omitted}")


class MDDriver:
    def __init__(self, action, integrator, trajectory_length, steps):
        self.action = action
        self.integrator = integrator
        self.trajectory_length = trajectory_length
        self.steps = steps


def md_driver(U, action, *, steps, trajectory_length=1.0, integrator=None):
    """molecular_dynamics.jl:425-483 validation and construction"""
    if len(U) != 4:
        raise ArgumentError(f"the gauge configuration has length {len(U)}; expected 4")
    if steps <= 0:
        raise ArgumentError(f"steps must be positive; got {steps}")
    if not math.isfinite(trajectory_length):
        raise ArgumentError(f"trajectory_length must be finite; got {trajectory_length}")
    if trajectory_length == 0:
        raise ArgumentError("trajectory_length must not be zero")
    if U.lattice.precision == "f32":
        tl32 = float(np.float32(trajectory_length))
        if tl32 == 0:
            raise ArgumentError("trajectory_length becomes zero after conversion to Float32")
        trajectory_length = tl32
    integrator = integrator if integrator is not None else QPQ()
    _validate_md_integrator(integrator, action)
    md_action_workspace(action, U)
    return MDDriver(action, integrator, trajectory_length, int(steps))


def md_step_size(driver):
    return driver.trajectory_length / driver.steps


def _configure(driver, U):
    return _configure_member(driver.action, U)


def md_hamiltonian(U, P, driver):
    """md_potential + p*p/2 (molecular_dynamics.jl:494-505)"""
    _check_same(U, P)
    if isinstance(driver.action, MDActionSet):
        return md_potential(driver.action, U) + 0.5 * U.lattice.momenta_dot()
    return U.lattice.hamiltonian(_configure(driver, U))


def update_gaugefields(U, P, step_size, driver):
    """U_mu <- exp(step_size P_mu) U_mu"""
    _check_same(U, P)
    if not math.isfinite(step_size):
        raise ArgumentError(f"the gauge-field step size must be finite; got {step_size}")
    U.lattice.exp_update(step_size)
    return U


def update_momenta(P, U, step_size, driver, group=None):
    """P <- P + step_size * force(U) (molecular_dynamics.jl:533-551); with an MDForceGroup, only
    the named members of the driver's MDActionSet (:553-583).  Each member's force and kick are
    one library call."""
    _check_same(U, P)
    if not math.isfinite(step_size):
        raise ArgumentError(f"the momentum step size must be finite; got {step_size}")
    for m in _providers(driver.action, group):
        U.lattice.force_kick(_configure_member(m, U), step_size)
    return P


def _md_fast_qpq(U, P, duration, nsteps, driver, fast):
    """n QPQ substeps of the fast group over `duration`, adjacent half link updates merged
    (_md_fast_qpq!, molecular_dynamics.jl:618-634)"""
    step_size = duration / nsteps
    update_gaugefields(U, P, step_size / 2, driver)
    for step in range(1, nsteps + 1):
        update_momenta(P, U, step_size, driver, fast)
        update_gaugefields(U, P, step_size / 2 if step == nsteps else step_size, driver)


def md_step(integrator, U, P, step_size, driver):
    if callable(integrator) and not isinstance(integrator, (QPQ, PQP, OmelyanPQP, SextonWeingarten)):
        return integrator(U, P, step_size, driver)
    if isinstance(integrator, SextonWeingarten):  # molecular_dynamics.jl:636-700
        if isinstance(integrator.ordering, QPQ):
            _md_fast_qpq(U, P, step_size / 2, integrator.n_fast, driver, integrator.fast)
            update_momenta(P, U, step_size, driver, integrator.slow)
            _md_fast_qpq(U, P, step_size / 2, integrator.n_fast, driver, integrator.fast)
        else:
            update_momenta(P, U, step_size / 2, driver, integrator.slow)
            _md_fast_qpq(U, P, step_size, integrator.n_fast, driver, integrator.fast)
            update_momenta(P, U, step_size / 2, driver, integrator.slow)
    elif isinstance(integrator, PQP):
        update_momenta(P, U, step_size / 2, driver)
        update_gaugefields(U, P, step_size, driver)
        update_momenta(P, U, step_size / 2, driver)
    elif isinstance(integrator, QPQ):
        update_gaugefields(U, P, step_size / 2, driver)
        update_momenta(P, U, step_size, driver)
        update_gaugefields(U, P, step_size / 2, driver)
    elif isinstance(integrator, OmelyanPQP):
        lam = integrator.lam
        update_momenta(P, U, lam * step_size, driver)
        update_gaugefields(U, P, 0.5 * step_size, driver)
        update_momenta(P, U, (1 - 2 * lam) * step_size, driver)
        update_gaugefields(U, P, 0.5 * step_size, driver)
        update_momenta(P, U, lam * step_size, driver)
    else:
        raise ArgumentError(f"md_step! is not implemented for {type(integrator).__name__}")


def md_trajectory(U, P, driver, *, diagnostics=True, fused=False):
    """One deterministic MD trajectory (no momentum refresh, no accept/reject).

    Built-in integrators run in the library's native driver (one C call per trajectory;
    ``fused=True`` merges adjacent half steps into single force + kick + exp kernels).
    Custom integrators (callables) are stepped from here."""
    _check_same(U, P)
    integ = driver.integrator
    if isinstance(integ, (QPQ, PQP, OmelyanPQP)) and not isinstance(driver.action, MDActionSet):
        coeff = _configure(driver, U)
        lam = integ.lam if isinstance(integ, OmelyanPQP) else OMELYAN_LAMBDA
        H = U.lattice.md_trajectory(coeff, driver.steps, driver.trajectory_length, integ.code, lam=lam,
                                    fused=fused, diagnostics=diagnostics)
        if not diagnostics:
            return None
        return {"initial_hamiltonian": H[0], "final_hamiltonian": H[1], "delta_hamiltonian": H[2]}
    h0 = md_hamiltonian(U, P, driver) if diagnostics else None
    eps = md_step_size(driver)
    for _ in range(driver.steps):
        md_step(integ, U, P, eps, driver)
    if not diagnostics:
        return None
    h1 = md_hamiltonian(U, P, driver)
    return {"initial_hamiltonian": h0, "final_hamiltonian": h1, "delta_hamiltonian": h1 - h0}


# ----------------------------------------------------------------------------- flow


class Gradientflow:
    def __init__(self, U, Nflow=1, eps=0.01):
        self.U = U
        self.Nflow = int(Nflow)
        self.eps = float(eps)


def gradient_flow(U, *, steps=1, step_size=0.01):
    if steps <= 0:
        raise ArgumentError(f"steps must be positive; got {steps}")
    if step_size <= 0:
        raise ArgumentError(f"step_size must be positive; got {step_size}")
    return Gradientflow(U, Nflow=steps, eps=step_size)


def flow(U, g, *, measure_every=0):
    """Nflow Luscher RK3 steps in place; with measure_every > 0 returns E(t) samples."""
    return U.lattice.flow(g.eps, g.Nflow, measure_every)


def energy_density(U):
    """E = 2 sum_{mu<nu} Re tr(1 - P_munu) / V (plaquette definition)"""
    return U.lattice.energy_density()


# ----------------------------------------------------------------------------- stout


class StoutSmearing:
    """A stack of isotropic plaquette stout layers (CovNeuralnet of STOUT_Layer)."""

    def __init__(self, U, levels, rho, mode="store"):
        if levels < 0:
            raise ArgumentError("levels must be nonnegative")
        if mode not in ("store", "recompute", "auto"):
            raise ArgumentError("mode must be 'store', 'recompute' or 'auto'")
        self.levels = int(levels)
        self.rho = float(rho)
        self.mode = mode

    def configure(self, U):
        U.lattice.stout_config(self.levels, self.rho, self.mode)


def stout_smearing(U, *, loops="plaquette", rho=0.1, levels=1, mode="auto"):
    """stout_smearing (API.jl:452-467): `levels` plaquette layers at the same rho"""
    names = [loops] if isinstance(loops, str) else list(loops)
    if names != ["plaquette"]:
        raise ArgumentError("only plaquette stout layers are implemented on the HIP backend")
    rhos = [rho] if np.isscalar(rho) else list(rho)
    if len(rhos) != 1:
        raise ArgumentError("rho must be a scalar or have one entry per loop")
    return StoutSmearing(U, levels, rhos[0], mode)


def smear(U, smearing, *, record=False):
    """smeared links (4, V, 9) complex128; record=True also returns the setting"""
    smearing.configure(U)
    out = U.lattice.stout_smear() if smearing.levels > 0 else U.to_host()
    return {"configuration": out, "smearing": smearing} if record else out


def back_prop(dSdU_smeared, smearing, U):
    """dS/dU_bare from dS/dU_smeared (both (4, V, 9) reference layout); re-smears U first"""
    smearing.configure(U)
    U.lattice.stout_smear(download=False)
    return U.lattice.stout_backward(dSdU_smeared)


class SmearedGaugeAction:
    """MD action provider: a GaugeAction evaluated on stout-smeared links
    (test/HMCstout_test_nowing.jl:36-118)."""

    def __init__(self, action: GaugeAction, smearing: StoutSmearing):
        self.action = action
        self.smearing = smearing

    @property
    def coefficient(self):
        return self.action.coefficient

    def configure(self, U):
        self.smearing.configure(U)

    def potential(self, U):
        self.configure(U)
        # the Hamiltonian smears the links itself (in FP64); minus the kinetic part
        return U.lattice.hamiltonian(self.coefficient) - 0.5 * U.lattice.momenta_dot()


# ----------------------------------------------------------------------------- ILDG files


class ILDG:
    """Header of an ILDG file (ILDG(filename), ildg_format.jl:686-695).

    ``len(ildg) == 1``; ``ildg[1]`` (1-based, as in the reference) is a dict with "L",
    "NC" and "precision".  A file without an ildg-format record reports L = None and the
    reference's default precision 64 (load_gaugefield!, ildg_format.jl:919-923)."""

    def __init__(self, filename):
        from .ildg import ildg_info

        self.filename = str(filename)
        info = ildg_info(filename)
        self.header = [{
            "L": list(info["L"]) if info["L"] else None,
            "NC": 3,
            "precision": info["precision"] or 64,
            "payload_bytes": info["payload_bytes"],
        }]

    def __len__(self):
        return len(self.header)

    def __getitem__(self, i):
        if not 1 <= i <= len(self.header):
            raise IndexError(f"ILDG index {i} out of range 1:{len(self.header)}")
        return self.header[i - 1]


def load_gaugefield(U, i, ildg, L, NC):
    """load_gaugefield!(U, i, ildg, L, NC) (ildg_format.jl:902-969): read record i into U."""
    L = [int(x) for x in L]
    if list(U.dims) != L:
        raise ArgumentError(f"lattice mismatch: U has {list(U.dims)} but L = {L}")
    if NC != U.NC:
        raise ArgumentError(f"NC mismatch: U has {U.NC} but NC = {NC}")
    ildg[i]  # range check
    U.lattice.load_ildg(ildg.filename)
    return U


def _ildg_precision(U, precision):
    if precision == "field":
        return 64 if U.eltype == np.complex128 else 32
    if precision not in (32, 64):
        raise ArgumentError(f"ILDG precision must be :field, 32, or 64, got {precision!r}")
    return int(precision)


def save_binarydata(U, filename, *, precision="field"):
    """save_binarydata(U, filename; precision) (ildg_format.jl:797-833): one ILDG file
    with an ildg-format record and the ildg-binary-data record."""
    U.lattice.save_ildg(filename, _ildg_precision(U, precision))


def save_configuration(filename, U, *, format="ildg", **kwargs):
    """save_configuration (API.jl:711-727); the HIP backend writes ILDG only."""
    if format == "ildg":
        return save_binarydata(U, filename, **kwargs)
    if format in ("jld2", "bridge"):
        raise ArgumentError(f"format :{format} is not provided by the HIP backend; use :ildg")
    raise ArgumentError(f"format must be :jld2, :bridge, or :ildg; got {format}")


def load_configuration_(U, filename, *, format="ildg"):
    """load_configuration!(U, filename; format) (API.jl:781-826); ILDG only here."""
    if format == "ildg":
        U.lattice.load_ildg(filename)
        return U
    if format in ("jld2", "bridge"):
        raise ArgumentError(f"format :{format} is not provided by the HIP backend; use :ildg")
    raise ArgumentError(f"format must be :jld2, :bridge, or :ildg; got {format}")


# ----------------------------------------------------------------------------- field strengths

_TOPO_PUBLIC = ("plaquette", "clover", "improved")


def topological_charge(U, *, method="plaquette"):
    """topological_charge(U; method) (AbstractGaugefields.jl:1452-1471)"""
    if method not in _TOPO_PUBLIC:
        raise ArgumentError("supported topological_charge methods are :plaquette, :clover, and :improved")
    return U.lattice.topological_charge(method)


def topological_charge_density(U, *, method="plaquette"):
    """topological_charge_density(U; method) (AbstractGaugefields.jl:1430-1446): an
    (NX, NY, NZ, NT) array indexed like the reference's q[ix, iy, iz, it]"""
    if method not in _TOPO_PUBLIC:
        raise ArgumentError("supported topological_charge_density methods are :plaquette, :clover, and :improved")
    return U.lattice.topological_charge_density(method).reshape(U.dims, order="F")


def calculate_energy_density(U):
    """clover energy density (calculate_energy_density, samples/measurements/energydensity.jl:68-76)"""
    return U.lattice.energy_density_clover()


# ----------------------------------------------------------------------------- heatbath


class Heatbath:
    """Heatbath(U, beta; ITERATION_MAX, seed, sweep, overrelaxation_sweep) (heatbathmodule.jl:50-97):
    checkerboard SU(3) heatbath / overrelaxation of the Wilson plaquette action at beta.
    Successful heatbath / overrelaxation calls advance sweep / overrelaxation_sweep."""

    def __init__(self, U, beta, *, ITERATION_MAX=10 ** 5, seed=0, sweep=0, overrelaxation_sweep=None):
        if ITERATION_MAX <= 0:
            raise ArgumentError("ITERATION_MAX must be positive")
        if any(int(d) % 2 for d in U.dims):
            raise ArgumentError(f"the checkerboard heatbath needs even lattice extents; got {list(U.dims)}")
        self.beta = float(beta)
        self.ITERATION_MAX = int(min(ITERATION_MAX, 2 ** 31 - 1))
        self.seed = int(seed)
        self.sweep = int(sweep)
        self.overrelaxation_sweep = int(sweep if overrelaxation_sweep is None else overrelaxation_sweep)
        self.failures = 0


def Heatbath_update(U, gauge_action, *, ITERATION_MAX=10 ** 5, seed=0, sweep=0, overrelaxation_sweep=None):
    """Heatbath_update(U, gauge_action; ...) (heatbathmodule.jl:148-200) for the plaquette action
    push!(action, beta/2, plaquette ∪ plaquette'): the same updater at beta = 2 * coefficient"""
    return Heatbath(U, 2.0 * gauge_action.coefficient, ITERATION_MAX=ITERATION_MAX, seed=seed, sweep=sweep,
                    overrelaxation_sweep=overrelaxation_sweep)


def heatbath(U, h):
    """heatbath!(U, h) (heatbathmodule.jl:2165-2181): one sweep, then h.sweep += 1.  Links whose
    SU(2) proposals were all rejected within ITERATION_MAX tries stay unchanged (h.failures)."""
    h.failures += U.lattice.heatbath(h.beta, h.seed, h.sweep, h.ITERATION_MAX)
    h.sweep += 1
    return U


def overrelaxation(U, h):
    """overrelaxation!(U, h): one microcanonical sweep, then h.overrelaxation_sweep += 1"""
    U.lattice.overrelaxation(h.seed, h.overrelaxation_sweep)
    h.overrelaxation_sweep += 1
    return U
