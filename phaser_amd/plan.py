"""Engine plans with the reference's field names and defaults (phaser/plan.py:43-173,
phaser/hooks/solver.py:79-98, phaser/types.py:125-167, phaser/hooks/schedule.py).

Plans are plain dataclasses; `from_dict` accepts the mapping form of a plan YAML engine
entry (`{'type': 'conventional', 'solver': 'lsqml', 'noise_model': 'amplitude', ...}`)."""
from dataclasses import dataclass, field, fields
import typing as t

import numpy


@dataclass(frozen=True)
class Flag:
    """Iteration flag (types.py:125-146): true when i > after, i < before and
    (i - after) % every == 0."""
    after: int = 0
    every: int = 1
    before: t.Optional[int] = None

    def __call__(self, args) -> bool:
        i = args['state'].iter.engine_iter
        return ((self.before is None or i < self.before) and i > self.after
                and (i - self.after) % self.every == 0)


def process_flag(flag):
    """bool or Flag -> callable(args) -> bool (types.py:149-160)."""
    if isinstance(flag, bool):
        return lambda _args: flag
    if isinstance(flag, dict):
        flag = Flag(**flag)
    return flag


def shuffle_groups_flag(props):
    """Whether groups are reshuffled each iteration: `shuffle_groups or not compact`
    (conventional/run.py:32, gradient/run.py:185), so an explicit `False` still shuffles
    sparse groups, and compact groups keep their order unless a flag says otherwise."""
    return process_flag(props.shuffle_groups or not props.compact)


def process_schedule(schedule):
    """float or callable(args) -> float (types.py:163-167); dict forms of the known
    schedules (constant / piecewise / expr, hooks/schedule.py:33-75) are accepted."""
    if isinstance(schedule, (int, float)):
        return lambda _args: float(schedule)
    if isinstance(schedule, dict):
        kind = schedule.get('type')
        if kind == 'constant':
            v = float(schedule['value'])
            return lambda _args: v
        if kind == 'piecewise':
            init = process_schedule(schedule['init'])
            steps = sorted(((int(k), process_schedule(v)) for k, v in schedule['steps'].items()), reverse=True)

            def piecewise(args):
                i = args['state'].iter.engine_iter
                for thresh, val in steps:
                    if i >= thresh:
                        return val(args)
                return init(args)
            return piecewise
        if kind == 'expr':
            expr = schedule['expr']
            return lambda args: float(eval(expr, {  # same evaluation namespace as schedule.py:60-71
                'i': args['state'].iter.engine_iter, 'iter': args['state'].iter,
                'state': args['state'], 'niter': args['niter'], 'np': numpy}))
        raise ValueError(f"Unknown schedule {schedule!r}")
    return schedule


@dataclass
class AmplitudeNoisePlan:
    gaussian_variance: float = 0.1
    eps: float = 1.0e-3
    offset: float = 0.0
    name: str = 'amplitude'


@dataclass
class AnscombeNoisePlan(AmplitudeNoisePlan):
    offset: float = 0.375
    name: str = 'anscombe'


@dataclass
class PoissonNoisePlan(AmplitudeNoisePlan):
    eps: float = 1.0e-3
    name: str = 'poisson'


NOISE_MODELS = {'amplitude': AmplitudeNoisePlan, 'anscombe': AnscombeNoisePlan, 'poisson': PoissonNoisePlan}


@dataclass
class LSQMLSolverPlan:
    stochastic: bool = True
    beta_object: t.Any = 1.0
    beta_probe: t.Any = 1.0
    illum_reg_object: t.Any = 1e-2
    illum_reg_probe: t.Any = 1e-2
    gamma: t.Any = 1e-4
    name: str = 'lsqml'


@dataclass
class EPIESolverPlan:
    beta_object: t.Any = 1.0
    beta_probe: t.Any = 1.0
    name: str = 'epie'


SOLVERS = {'lsqml': LSQMLSolverPlan, 'epie': EPIESolverPlan}


@dataclass
class SteepestDescentPositionSolverPlan:
    step_size: float = 1e-2
    max_step_size: t.Optional[float] = None
    name: str = 'steepest_descent'


@dataclass
class MomentumPositionSolverPlan:
    step_size: float = 1e-2
    max_step_size: t.Optional[float] = None
    momentum: float = 0.9
    name: str = 'momentum'


POSITION_SOLVERS = {'steepest_descent': SteepestDescentPositionSolverPlan, 'momentum': MomentumPositionSolverPlan}


# -------------------------------------------- constraints and cost regularizers
# (phaser/hooks/regularization.py:36-118: names, fields and defaults)
@dataclass
class ClampObjectAmplitudePlan:
    amplitude: float = 1.1
    name: str = 'clamp_object_amplitude'


@dataclass
class LimitProbeSupportPlan:
    max_angle: float = None
    name: str = 'limit_probe_support'

    def __post_init__(self):
        if self.max_angle is None:
            raise ValueError("limit_probe_support needs 'max_angle' (mrad)")


@dataclass
class RegularizeLayersPlan:
    sigma: float = 50.0
    weight: float = 0.9
    name: str = 'layers'


@dataclass
class ObjLowPassPlan:
    max_freq: float = 0.4
    name: str = 'obj_low_pass'


@dataclass
class ObjGaussianPlan:
    sigma: float = None
    weight: float = 0.9
    name: str = 'obj_gaussian'

    def __post_init__(self):
        if self.sigma is None:
            raise ValueError("obj_gaussian needs 'sigma' (angstrom)")


@dataclass
class CostRegularizerPlan:
    cost: float = None
    eps: float = 1.0e-8
    name: str = ''

    def __post_init__(self):
        if self.cost is None:
            raise ValueError(f"cost regularizer {self.name!r} needs 'cost'")


@dataclass
class UnsupportedPlan:
    """A constraint or regularizer phaser knows but this backend does not run; engines
    raise NotImplementedError naming it."""
    name: str = ''
    props: t.Any = None


GROUP_CONSTRAINTS = {'clamp_object_amplitude': ClampObjectAmplitudePlan, 'limit_probe_support': LimitProbeSupportPlan,
                     'obj_low_pass': ObjLowPassPlan, 'obj_gaussian': ObjGaussianPlan}
ITER_CONSTRAINTS = dict(GROUP_CONSTRAINTS, layers=RegularizeLayersPlan)
REGULARIZER_NAMES = {'obj_l1': 'obj_l1', 'obj_l2': 'obj_l2', 'obj_tikh': 'obj_tikh', 'obj_tikhonov': 'obj_tikh',
                     'layers_tv': 'layers_tv', 'layers_tikh': 'layers_tikh', 'layers_tikhonov': 'layers_tikh'}
# known to phaser (hooks/regularization.py:76-118), not run by this backend
UNSUPPORTED = {'opr_gaussian', 'tilt_gaussian', 'remove_phase_ramp', 'obj_phase_l1', 'obj_recip_l1', 'obj_tv',
               'probe_phase_tikh', 'probe_phase_tikhonov', 'probe_recip_tv', 'probe_recip_tikh',
               'probe_recip_tikhonov'}


def _resolve_item(val, known, what):
    if not isinstance(val, (str, dict)):
        return val
    d = {'type': val} if isinstance(val, str) else dict(val)
    kind = d.get('type')
    if what == 'regularizer' and kind in REGULARIZER_NAMES:
        d.pop('type')
        bad = set(d) - {'cost', 'eps'}
        if bad:
            raise ValueError(f"Unknown {what} options {sorted(bad)}")
        return CostRegularizerPlan(name=REGULARIZER_NAMES[kind], **d)
    if kind in UNSUPPORTED:
        d.pop('type')
        return UnsupportedPlan(kind, d)
    return _resolve(d, known, what)


def _resolve_list(vals, known, what):
    return [_resolve_item(v, known, what) for v in (vals or [])]


@dataclass
class EnginePlan:
    """phaser/plan.py:43-83"""
    sim_shape: t.Optional[t.Tuple[int, int]] = None
    resize_method: str = 'pad_crop'
    probe_modes: int = 1
    base_mode_power: float = 0.7
    bwlim_frac: t.Optional[float] = 2 / 3
    obj_pad_px: float = 5.0
    slices: t.Any = None
    niter: int = 10
    grouping: t.Optional[int] = None
    compact: bool = False
    shuffle_groups: t.Any = None
    buffer_n_groups: int = 2
    update_probe: t.Any = True
    update_object: t.Any = True
    update_positions: t.Any = False
    update_tilt: t.Any = False
    calc_error: t.Any = field(default_factory=lambda: Flag(every=1))
    calc_error_fraction: float = 0.1
    save: t.Any = False
    save_images: t.Any = False
    early_termination: t.Optional[int] = None
    early_termination_smoothing: float = 0.9
    send_every_group: bool = False
    save_options: t.Any = None


@dataclass
class ConventionalEnginePlan(EnginePlan):
    """phaser/plan.py:126-132."""
    noise_model: t.Any = None
    solver: t.Any = None
    position_solver: t.Any = None
    group_constraints: t.List[t.Any] = field(default_factory=list)
    iter_constraints: t.List[t.Any] = field(default_factory=list)

    def __post_init__(self):
        self.noise_model = _resolve(self.noise_model, NOISE_MODELS, 'noise_model')
        self.solver = _resolve(self.solver, SOLVERS, 'solver')
        if self.position_solver is not None:
            self.position_solver = _resolve(self.position_solver, POSITION_SOLVERS, 'position_solver')
        self.group_constraints = _resolve_list(self.group_constraints, GROUP_CONSTRAINTS, 'group constraint')
        self.iter_constraints = _resolve_list(self.iter_constraints, ITER_CONSTRAINTS, 'iter constraint')


@dataclass
class SGDSolverPlan:
    learning_rate: t.Any = 1.0
    momentum: t.Any = None
    nesterov: bool = True
    name: str = 'sgd'


@dataclass
class AdamSolverPlan:
    learning_rate: t.Any = 1.0
    b1: float = 0.9
    b2: float = 0.999
    eps: float = 1.0e-8
    eps_root: float = 0.0
    nesterov: bool = False
    name: str = 'adam'


@dataclass
class PolyakSGDSolverPlan:
    max_learning_rate: t.Any = 1.0
    f_min: float = 0.0
    scaling: t.Any = 1.0
    eps: float = 0.0
    name: str = 'polyak_sgd'


GRADIENT_SOLVERS = {'sgd': SGDSolverPlan, 'adam': AdamSolverPlan, 'polyak_sgd': PolyakSGDSolverPlan}


@dataclass
class GradientEnginePlan(EnginePlan):
    """phaser/plan.py:135-141.  `solvers` maps a frozenset of variable names
    ('object', 'probe', 'positions', 'tilt') to a solver plan."""
    noise_model: t.Any = None
    solvers: t.Dict[t.Any, t.Any] = field(default_factory=dict)
    regularizers: t.List[t.Any] = field(default_factory=list)
    group_constraints: t.List[t.Any] = field(default_factory=list)
    iter_constraints: t.List[t.Any] = field(default_factory=list)

    def __post_init__(self):
        self.noise_model = _resolve(self.noise_model, NOISE_MODELS, 'noise_model')
        solvers = {}
        for k, v in self.solvers.items():
            key = frozenset([k]) if isinstance(k, str) else frozenset(k)
            solvers[key] = _resolve(v, GRADIENT_SOLVERS, 'solver')
        self.solvers = solvers
        self.regularizers = _resolve_list(self.regularizers, {}, 'regularizer')
        self.group_constraints = _resolve_list(self.group_constraints, GROUP_CONSTRAINTS, 'group constraint')
        self.iter_constraints = _resolve_list(self.iter_constraints, ITER_CONSTRAINTS, 'iter constraint')


def _resolve(val, known, what):
    if val is None:
        raise ValueError(f"'{what}' must be specified")
    if isinstance(val, str):
        if val not in known:
            raise ValueError(f"Unknown {what} '{val}'")
        return known[val]()
    if isinstance(val, dict):
        d = dict(val)
        kind = d.pop('type')
        if kind not in known:
            raise ValueError(f"Unknown {what} '{kind}'")
        cls = known[kind]
        names = {f.name for f in fields(cls)}
        bad = set(d) - names
        if bad:
            raise ValueError(f"Unknown {what} options {sorted(bad)}")
        return cls(**d)
    return val


def from_dict(d: t.Mapping[str, t.Any]):
    """Engine plan from the mapping form of a plan YAML `engines:` entry."""
    d = dict(d)
    kind = d.pop('type', 'conventional')
    if 'calc_error' in d and isinstance(d['calc_error'], dict):
        d['calc_error'] = Flag(**d['calc_error'])
    cls = {'conventional': ConventionalEnginePlan, 'gradient': GradientEnginePlan}[kind]
    return cls(**d)


# ------------------------------------------------------------------ reconstruction plans
@dataclass
class SaveOptions:
    """phaser/plan.py:29-40."""
    images: t.Tuple[str, ...] = ('probe', 'object_phase_stack')
    crop_roi: bool = True
    unwrap_phase: bool = False
    img_dtype: str = '16bit'
    plot_ext: str = 'svg'
    plot_dpi: int = 300
    out_dir: str = '{name}'
    img_fmt: str = '{type}_iter{iter.total_iter:03}.{ext}'
    hdf5_fmt: str = 'iter{iter.total_iter:03}.h5'


def slice_thicknesses(slices):
    """The thickness list of a plan's `slices` entry (types.py:95-121): {thicknesses},
    {n, slice_thickness} or {n, total_thickness}."""
    if slices is None:
        return None
    if isinstance(slices, (list, tuple)):
        return [float(v) for v in slices]
    s = dict(slices)
    if 'thicknesses' in s:
        return [float(v) for v in s['thicknesses']]
    n = int(s['n'])
    if 'slice_thickness' in s:
        return [float(s['slice_thickness'])] * n
    if 'total_thickness' in s:
        return [float(s['total_thickness']) / n] * n
    raise ValueError(f"Invalid slices {slices!r}")


@dataclass
class ReconsPlan:
    """phaser/plan.py:176-198: a whole reconstruction (data, initial state, engines)."""
    name: str = None
    raw_data: t.Any = None
    engines: t.List[t.Any] = field(default_factory=list)
    file_type: str = 'phaser_plan'
    version: str = '1.0'
    backend: t.Optional[str] = None
    dtype: str = 'float32'
    wavelength: t.Optional[float] = None
    post_load: t.List[t.Any] = field(default_factory=list)
    init: t.Dict[str, t.Any] = field(default_factory=dict)
    post_init: t.List[t.Any] = field(default_factory=list)
    slices: t.Any = None

    def __post_init__(self):
        if self.file_type != 'phaser_plan':
            raise ValueError(f"Expected file_type 'phaser_plan', got {self.file_type!r}")
        if str(self.version) != '1.0':
            raise ValueError(f"Unsupported plan version {self.version!r} (expected '1.0')")
        if self.name is None or self.raw_data is None:
            raise ValueError("A plan needs 'name' and 'raw_data'")
        if self.dtype not in ('float32', 'float64'):
            raise ValueError(f"Invalid dtype {self.dtype!r}")
        if self.backend not in (None, 'numpy', 'cupy', 'jax', 'mi355x'):
            raise ValueError(f"Unknown backend {self.backend!r}")
        bad = set(self.init or {}) - {'state', 'scan', 'tilt', 'probe', 'object'}
        if bad:
            raise ValueError(f"Unknown init keys {sorted(bad)}")
        self.init = dict(self.init or {})
        self.slices = slice_thicknesses(self.slices)
        self.engines = [e if isinstance(e, EnginePlan) else from_dict(_engine_entry(e)) for e in self.engines]


def _engine_entry(e):
    d = dict(e)
    if isinstance(d.get('slices'), (dict, list)):
        d['slices'] = slice_thicknesses(d['slices'])
    if isinstance(d.get('save_options'), dict):
        d['save_options'] = SaveOptions(**d['save_options'])
    for k in ('shuffle_groups', 'update_probe', 'update_object', 'update_positions', 'update_tilt', 'save',
              'save_images'):
        if isinstance(d.get(k), dict):
            d[k] = Flag(**d[k])
    if isinstance(d.get('sim_shape'), list):
        d['sim_shape'] = tuple(d['sim_shape'])
    return d


def plans_from_yaml(path):
    """Every plan (YAML document) in a plan file (ReconsPlan.from_yaml_all, main.py:17)."""
    import yaml
    with open(path) as f:
        docs = [d for d in yaml.safe_load_all(f) if d is not None]
    return [ReconsPlan(**d) for d in docs]
