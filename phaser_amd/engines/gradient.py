"""Gradient-descent engine on the MI355X backend, with a hand-written adjoint.

`run_engine(args, props)` follows phaser/engines/gradient/run.py:156-290: seeded groups,
dry-run probe rescale (:196-209), per-group loss + gradient of the multislice model
(:293-349), the steepest-descent direction -conj(g_jax) scaled by 1/B for the probe and
1/(probe_int B) for the other variables (:324-330), group solvers for object/probe,
per-iteration solvers for positions (:266-274), loss history in ProgressState.

The model's loss and gradient come from HIP kernels (phx_farfield_gradient,
phx_gradient_backward, phx_gradient_probe_finish); `MultisliceLoss` exposes them as a
torch.autograd.Function.  jax's complex gradient is the conjugate of torch's, so the
descent direction -conj(g_jax) equals -g_torch, which is what the kernels consume.
With `recompute=True` only psi_0 and two ping-pong slices are stored and each slice's
wave is recomputed from psi_0 for the adjoint (memory S -> 3 planes per position/mode,
work +S^2/2 slice steps).
"""
import ctypes
import logging

import numpy

from .. import _lib
from ..constraints import CostRegularizers, make_constraints
from ..device import FUSED_MAX_N, SPLIT_N, GroupWorkspace, host_patterns, stream_handle, INT32_MAX
from ..groups import GroupManager
from ..optics import make_propagators
from ..plan import process_flag, process_schedule, shuffle_groups_flag
from .conventional import Observer

logger = logging.getLogger(__name__)

PER_ITER_VARS = frozenset({'positions', 'tilt'})
# the object-gradient gather of group g also clears group g+1's accumulators
# (phx_gradient_object_gather_begin) instead of a separate phx_group_begin
FOLD_GROUP_BEGIN = True


def red_layout(S, B, M):
    """Offsets of the FP64 reduction block (mirrors red_layout in phx_engine.hip)."""
    o = {}
    k = 0
    for name, size in (('num_obj', S * B), ('den_obj', S), ('numP', B * M), ('denP', 1), ('norm', B),
                       ('pos_num', 2 * B), ('pos_den', 2), ('loss', 1)):
        o[name] = k
        k += size
    o['total'] = k
    return o


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class MultisliceModel:
    """Per-group loss and gradient of the multislice model on a GroupWorkspace."""

    def __init__(self, ws: GroupWorkspace, poisson: bool, fused=None):
        self.ws = ws
        self.poisson = int(poisson)
        self.L = red_layout(ws.S, ws.B, ws.M)
        self._pending = None
        self._farfield_pending = False   # far-field step deferred to the adjoint sweep
        self._begun = None               # group slot whose accumulators a gather cleared
        # whole-sweep kernels need every slice's wave stored (not the recompute mode) and,
        # with several slices, the per-slice wave-gradient planes (chi_slices)
        ok = ws.psi_slots >= ws.S and (ws.S == 1 or ws.chi_slices is not None)
        self.fused = ok if fused is None else (bool(fused) and ok)
        if ws.n > FUSED_MAX_N:
            self.fused = False   # per-slice kernels for wide planes (see GroupWorkspace)
        # split planes: the forward model as one launch sequence (phx_group_forward folds
        # each slice step's last row pass into the next step's first); the adjoint per slice
        self.split_forward = ws.n == SPLIT_N and ws.psi_slots >= ws.S
        # per-slice adjoint with every slice's wave gradient kept (chi_slices): the object
        # gradient is gathered per tile afterwards, as in the fused path, not scattered
        self.gather = ws.psi_slots >= ws.S and (ws.S == 1 or ws.chi_slices is not None)
        # with the far-field step in the adjoint sweep, the forward sweep is launched with
        # it (phx_gradient_group_sweeps schedules both as plane tasks when the group has
        # more planes than the device has CUs)
        self.group_sweeps = True
        self._forward_pending = False

    def dry_run(self, g):
        """errors[b] = sum(D) / model intensity of every slot (gradient/run.py:408-438)."""
        ws, d, st = self.ws, self.ws.desc_refs[g], stream_handle()
        _lib.call("phx_group_begin", d, st)
        self._begun = None
        ws.fill_tilt(g, st)
        _lib.call("phx_probe_spectra", d, st)
        if self.fused or self.split_forward:
            # the whole forward model in one launch; its stored intensity gives the sums
            _lib.call("phx_group_forward", d, st)
            _lib.call("phx_dry_run_stats_intensity", d, st)
            return
        _lib.call("phx_shift_probes", d, 0, int(ws.S > 1), st)
        for i in range(1, ws.S - 1):
            _lib.call("phx_forward_slice", d, i, 0, st)
        _lib.call("phx_dry_run_stats_gradient", d, st)

    def forward(self, g, probe_spectra=True, loss_in_backward=False):
        """Launch the forward model and the loss gradient w.r.t. the exit wave; the group
        loss lands in ws.red[g, loss].  probe_spectra=False reuses the probe spectra of
        the previous forward (valid while the probe is unchanged).  loss_in_backward=True
        (fused path) leaves the far-field gradient step and the loss to the adjoint sweep
        of backward(), which computes them in its first pass, and with group_sweeps the
        forward sweep too, which backward() launches with the adjoint."""
        ws, d, st = self.ws, self.ws.desc_refs[g], stream_handle()
        if self._begun != g:
            _lib.call("phx_group_begin", d, st)
        self._begun = None
        if probe_spectra or not self.fused:
            _lib.call("phx_probe_spectra", d, st)
        # backward() must follow forward() of the same group, so the adjoint finds this
        # group's tilted propagators in place
        ws.fill_tilt(g, st)
        self._farfield_pending = False
        self._forward_pending = False
        if self.fused:
            # every slice in one launch, then the far-field gradient and loss
            if loss_in_backward and self.group_sweeps:
                self._forward_pending = True
            else:
                _lib.call("phx_group_forward", d, st)
            if loss_in_backward:
                self._farfield_pending = True
            else:
                _lib.call("phx_farfield_gradient_update", d, self.poisson, st)
        elif self.split_forward:
            _lib.call("phx_group_forward", d, st)
            _lib.call("phx_farfield_gradient_update", d, self.poisson, st)
        else:
            _lib.call("phx_shift_probes", d, 0, int(ws.S > 1), st)
            for i in range(1, ws.S - 1):
                _lib.call("phx_forward_slice", d, i, 0, st)
            _lib.call("phx_farfield_gradient", d, self.poisson, st)
        self._pending = g

    def loss_tensor(self, g):
        return self.ws.red[g, self.L['loss']]

    def backward(self, g, want_object=True, want_probe=True, want_positions=True, next_group=None):
        """Adjoint pass for the group whose forward ran last; gradients land in
        ws.grad_obj (S, H, W), ws.grad_probe (M, n*n) and ws.red[g, pos_num].  With
        `next_group` (the slot whose forward comes next), the object-gradient gather also
        clears that slot's accumulators, and its forward skips phx_group_begin."""
        if self._pending != g:
            raise RuntimeError("backward() must follow forward() of the same group")
        ws, d, st = self.ws, self.ws.desc_refs[g], stream_handle()
        if self.fused:
            # the sweep stores each slice's wave gradient; one gather per object tile then
            # forms every slice's object gradient (all pixels written, no atomics)
            if self._forward_pending:
                _lib.call("phx_gradient_group_sweeps", d, self.poisson, int(want_object), int(want_probe),
                          int(want_positions), st)
                self._forward_pending = self._farfield_pending = False
            elif self._farfield_pending:
                _lib.call("phx_gradient_sweep_farfield", d, self.poisson, int(want_object), int(want_probe),
                          int(want_positions), st)
                self._farfield_pending = False
            else:
                _lib.call("phx_gradient_sweep", d, int(want_object), int(want_probe), int(want_positions), st)
            if want_object and next_group is not None and next_group != g:
                _lib.call("phx_gradient_object_gather_begin", d, 0, ws.S, ws.desc_refs[next_group], st)
                self._begun = next_group
            elif want_object:
                _lib.call("phx_gradient_object_gather", d, 0, ws.S, st)
            if want_probe:
                _lib.call("phx_gradient_probe_finish", d, st)
            self._pending = None
            return
        recompute = ws.psi_slots < ws.S
        scatter = want_object and not self.gather
        if scatter:
            _lib.call("phx_fill_u32", _p(ws.grad_obj), 0, ws.grad_obj.numel() * 2, st)
        for i in range(ws.S - 1, -1, -1):
            if recompute and 1 <= i <= ws.S - 3:
                for j in range(0, i):
                    _lib.call("phx_forward_slice", d, j, 0, st)
            _lib.call("phx_gradient_backward", d, i, int(scatter), int(want_probe), int(want_positions), st)
        if want_object and self.gather:
            if next_group is not None and next_group != g:
                _lib.call("phx_gradient_object_gather_begin", d, 0, ws.S, ws.desc_refs[next_group], st)
                self._begun = next_group
            else:
                _lib.call("phx_gradient_object_gather", d, 0, ws.S, st)
        if want_probe:
            _lib.call("phx_gradient_probe_finish", d, st)
        self._pending = None

    def position_grads(self, g):
        """(B, 2) dL/dpos tensor view for group slot g."""
        o = self.L['pos_num']
        return self.ws.red[g, o:o + 2 * self.ws.B].view(self.ws.B, 2)


def _check_plane_side(state, recompute):
    if int(state.probe.data.shape[-1]) == SPLIT_N and recompute:
        raise NotImplementedError(f"{SPLIT_N} x {SPLIT_N} planes: the MI355X gradient engine stores every slice's "
                                  "wave at this size (no recompute mode)")


def make_model(state, patterns, pattern_mask, noise, *, groups, bwlim_frac=2 / 3, recompute=False):
    """MultisliceModel over `state` (moved to the GPU if needed) for the given groups
    (list of index arrays into the scan shape).  Returns (model, device state)."""
    if not hasattr(state.object.data, 'is_cuda'):
        state = state.to_device()
    _check_plane_side(state, recompute)
    props = make_propagators(state.probe.sampling, state.wavelength, state.object.thicknesses, bwlim_frac)
    bmax = max(numpy.asarray(g).shape[-1] for g in groups)
    ws = GroupWorkspace(state, host_patterns(patterns), numpy.asarray(pattern_mask), props,
                        noise_offset=noise.offset, noise_eps=noise.eps, noise_var=1.0 + noise.gaussian_variance,
                        n_groups=len(groups), bmax=bmax, psi_slots=3 if recompute else 0, gradients=True,
                        keep_chi=not recompute)
    ws.set_positions(state.object.sampling, state.scan)
    ws.set_groups(list(groups))
    return MultisliceModel(ws, poisson=(noise.name == 'poisson')), state


def _autograd_function():
    import torch

    class MultisliceLoss(torch.autograd.Function):
        """loss = L(object, probe, positions) of one group, with the hand-written adjoint.

        `obj` and `probe` must be the tensors the model's workspace was built on; the
        gradients are in torch's convention (for a real loss L of complex z, 2 dL/dconj(z))."""

        @staticmethod
        def forward(ctx, obj, probe, positions, model, g):
            ws = model.ws
            if obj.data_ptr() != ws.object.data_ptr() or probe.data_ptr() != ws.probe.data_ptr():
                raise ValueError("MultisliceLoss: object/probe are not the workspace's tensors")
            model.forward(g)
            ctx.model, ctx.g, ctx.pos_shape = model, g, positions.shape
            return model.loss_tensor(g).to(torch.float32).clone()

        @staticmethod
        def backward(ctx, grad_out):
            model, g = ctx.model, ctx.g
            ws = model.ws
            model.backward(g)
            scale = grad_out.to(torch.float32)
            go = ws.grad_obj.clone() * scale
            gp = ws.grad_probe.view(ws.probe.shape).clone() * scale
            gpos = torch.zeros((ws.P, 2), dtype=torch.float32, device=ws.grad_obj.device)
            n = int(ws.sizes_host[g])
            idx = torch.from_numpy(ws.table_host[g, :n].astype(numpy.int64)).to(gpos.device)
            gpos[idx] = model.position_grads(g)[:n].to(torch.float32) * scale
            return go, gp, gpos.view(ctx.pos_shape), None, None

    return MultisliceLoss


_MLF = None


def multislice_loss(obj, probe, positions, model, g):
    """torch.autograd entry point: MultisliceLoss.apply(obj, probe, positions, model, g)."""
    global _MLF
    if _MLF is None:
        _MLF = _autograd_function()
    return _MLF.apply(obj, probe, positions, model, g)


# ------------------------------------------------------------------------ solvers
class _Solver:
    """optax-equivalent transform on the steepest-descent direction (solvers.py:48-138)."""

    def __init__(self, plan, params):
        self.plan = plan
        self.params = frozenset(params)
        self.name = plan.name
        self.hp = {}
        self.state = {}
        self.count = 0

    def _hparams(self):
        if self.name == 'sgd':
            names = ['learning_rate'] + (['momentum'] if self.plan.momentum is not None else [])
        elif self.name == 'adam':
            names = ['learning_rate']
        else:
            names = ['max_learning_rate', 'scaling']
        return names

    def update_for_iter(self, state, niter):
        for k in self._hparams():
            self.hp[k] = float(process_schedule(getattr(self.plan, k))({'state': state, 'niter': niter}))

    # device parameters -------------------------------------------------------------
    def init_device(self, name, tensor):
        import torch
        if self.name == 'sgd' and self.plan.momentum is not None:
            self.state[name] = (torch.zeros_like(tensor),)
        elif self.name == 'adam':
            self.state[name] = (torch.zeros_like(tensor), torch.zeros(tensor.shape, dtype=torch.float32,
                                                                      device=tensor.device))
        else:
            self.state[name] = ()

    def step_device(self, items, loss_ptr, sqnorm, flag, flag_value):
        """items: [(name, param tensor, grad tensor, scale)]."""
        st = stream_handle()
        self.count += 1
        if self.name == 'polyak_sgd':
            _lib.call("phx_fill_u32", _p(sqnorm), 0, 2, st)
            for _n, _prm, grd, sc in items:
                _lib.call("phx_sqnorm_accumulate", _p(grd), grd.numel(), sc, _p(sqnorm), st)
        for name, prm, grd, sc in items:
            n = prm.numel()
            if self.name == 'sgd':
                mom = self.plan.momentum is not None
                tr = self.state[name][0] if mom else None
                _lib.call("phx_sgd_step", _p(prm), _p(grd), None if tr is None else _p(tr), n, sc,
                          self.hp['learning_rate'], self.hp.get('momentum', 0.0), int(self.plan.nesterov), int(mom),
                          _p(flag), flag_value, st)
            elif self.name == 'adam':
                mu, nu = self.state[name]
                _lib.call("phx_adam_step", _p(prm), _p(grd), _p(mu), _p(nu), n, sc, self.hp['learning_rate'],
                          self.plan.b1, self.plan.b2, self.plan.eps, self.plan.eps_root, self.count,
                          int(self.plan.nesterov), _p(flag), flag_value, st)
            else:
                _lib.call("phx_polyak_step", _p(prm), _p(grd), n, sc, self.hp['scaling'],
                          self.hp['max_learning_rate'], self.plan.f_min, self.plan.eps, _p(sqnorm), loss_ptr,
                          _p(flag), flag_value, st)

    # host parameters (per-iteration variables: positions) ------------------------------
    def step_host(self, name, grad, scale, loss):
        """Real-valued update of a per-iteration variable (numpy); returns the update."""
        dv = -grad / scale
        self.count += 1
        if self.name == 'sgd':
            if self.plan.momentum is not None:
                t = dv + self.hp['momentum'] * self.state.get(name, numpy.zeros_like(dv))
                self.state[name] = t
                u = dv + self.hp['momentum'] * t if self.plan.nesterov else t
            else:
                u = dv
            return self.hp['learning_rate'] * u
        if self.name == 'adam':
            mu, nu = self.state.get(name, (numpy.zeros_like(dv), numpy.zeros_like(dv)))
            b1, b2 = self.plan.b1, self.plan.b2
            mu = (1 - b1) * dv + b1 * mu
            nu = (1 - b2) * dv**2 + b2 * nu
            self.state[name] = (mu, nu)
            c = self.count
            if self.plan.nesterov:
                mh = b1 * mu / (1 - b1**(c + 1)) + (1 - b1) * dv / (1 - b1**c)
            else:
                mh = mu / (1 - b1**c)
            u = mh / (numpy.sqrt(nu / (1 - b2**c) + self.plan.eps_root) + self.plan.eps)
            return self.hp['learning_rate'] * u
        u1 = self.hp['scaling'] * dv
        sq = float(numpy.sum(u1**2))
        step = 0.0 if sq + self.plan.eps <= numpy.finfo(float).eps else \
            min((loss - self.plan.f_min) / (sq + self.plan.eps), self.hp['max_learning_rate'])
        return step * u1


def process_solvers(props):
    """Split solvers into per-group and per-iteration ones (gradient/run.py:30-66)."""
    seen, dup = set(), set()
    group_solvers, iter_solvers = [], []
    for vars_, plan in props.solvers.items():
        if not vars_:
            continue
        dup |= vars_ & seen
        seen |= vars_
        if vars_ <= PER_ITER_VARS:
            iter_solvers.append(_Solver(plan, vars_))
            continue
        if vars_ & PER_ITER_VARS:
            raise ValueError("The same solver can't handle both per-iteration and per-group variables")
        group_solvers.append(_Solver(plan, vars_))
    if dup:
        raise ValueError(f"Duplicate solvers for variable(s) {', '.join(map(repr, sorted(dup)))}.")
    return frozenset(seen), group_solvers, iter_solvers


def run_engine(args, props, *, recompute=False):
    """Gradient engine (phaser/engines/gradient/run.py:156-290) on the HIP backend."""
    import torch
    observer = args.get('observer') or Observer()
    state = args['state']
    seed = args.get('seed')
    if not hasattr(state.object.data, 'is_cuda'):
        state = state.to_device()
    _check_plane_side(state, recompute)
    noise = props.noise_model
    all_vars, group_solvers, iter_solvers = process_solvers(props)
    if 'tilt' in all_vars:
        # fixed per-position tilts run (tilted propagators); their gradient does not
        raise NotImplementedError("tilt is not a variable of the MI355X gradient engine")

    flags = {'probe': process_flag(props.update_probe), 'object': process_flag(props.update_object),
             'positions': process_flag(props.update_positions), 'tilt': process_flag(props.update_tilt)}
    shuffle_groups = shuffle_groups_flag(props)
    groups = GroupManager(state.scan, props.grouping, props.compact, seed)
    patterns = host_patterns(args['data'].patterns)   # numpy, or float32 already on the GPU
    mask = numpy.asarray(args['data'].pattern_mask)

    observer.init_engine(state, recons_name=args.get('recons_name', ''), plan=props, noise_model=noise.name)
    start_i = state.iter.total_iter
    props_arr = make_propagators(state.probe.sampling, state.wavelength, state.object.thicknesses, props.bwlim_frac)
    n_pos = int(numpy.prod(state.scan.shape[:-1]))
    G = len(groups)
    B = int(numpy.ceil(n_pos / G))
    ws = GroupWorkspace(state, patterns, mask, props_arr, noise_offset=noise.offset, noise_eps=noise.eps,
                        noise_var=1.0 + noise.gaussian_variance, n_groups=G, bmax=B,
                        psi_slots=3 if recompute else 0, gradients=True, keep_chi=not recompute)
    ws.set_positions(state.object.sampling, state.scan)
    model = MultisliceModel(ws, poisson=(noise.name == 'poisson'))
    regularizers = CostRegularizers(props.regularizers, ws.S)
    group_constraints = make_constraints(props.group_constraints, ws, state, 'group constraint')
    iter_constraints = make_constraints(props.iter_constraints, ws, state, 'iter constraint')

    # dry run: rescale the probe (gradient/run.py:196-209)
    glist = list(groups.iter(state.scan))
    ws.set_groups(glist)
    for g in range(G):
        model.dry_run(g)
    errs = ws.errors.cpu().numpy()
    factors = numpy.concatenate([errs[g, :ws.sizes_host[g]] for g in range(G)]).astype(numpy.float32)
    rescale = float(numpy.mean(factors))
    ws.scale_probe(numpy.sqrt(rescale))
    probe_int = float(torch.sum(torch.abs(ws.probe)**2).item())
    observer.start_engine(state)

    for s in group_solvers:
        for name in s.params:
            s.init_device(name, ws.object if name == 'object' else ws.probe)
    sqnorm = torch.zeros(1, dtype=torch.float64, device=ws.object.device)

    for i in range(1, props.niter + 1):
        state.iter.engine_iter = i
        state.iter.total_iter = start_i + i
        fargs = {'state': state, 'niter': props.niter}
        iter_vars = all_vars & {k for k, f in flags.items() if f(fargs)}
        for s in group_solvers + iter_solvers:
            s.update_for_iter(state, props.niter)
        glist = list(groups.iter(state.scan, i, shuffle_groups(fargs)))
        ws.set_groups(glist)
        ws.reset_finite_flag()
        ws.check_state_finite(0)
        want_obj, want_probe, want_pos = ('object' in iter_vars), ('probe' in iter_vars), ('positions' in iter_vars)
        for g in range(G):
            # the probe spectra change only with the probe
            model.forward(g, probe_spectra=(g == 0 or want_probe), loss_in_backward=True)
            model.backward(g, want_object=want_obj, want_probe=want_probe, want_positions=want_pos,
                           next_group=(g + 1 if FOLD_GROUP_BEGIN and g + 1 < G else None))
            bsz = float(ws.sizes_host[g])
            loss_ptr = ctypes.c_void_p(ws.red[g].data_ptr() + 8 * model.L['loss'])
            # cost regularizers on the pre-update object (run.py:399-403): their loss joins
            # the group loss, their gradient the object's
            regularizers.apply(ws, int(ws.sizes_host[g]), loss_ptr, want_obj)
            for s in group_solvers:
                items = []
                if 'object' in s.params and want_obj:
                    items.append(('object', ws.object, ws.grad_obj, probe_int * bsz))
                if 'probe' in s.params and want_probe:
                    items.append(('probe', ws.probe, ws.grad_probe.view(ws.probe.shape), bsz))
                if items:
                    s.step_device(items, loss_ptr, sqnorm, ws.finite_flag, g)
            # group constraints after the solver steps; the reference's finite check
            # follows them (run.py:344-347, :261), so a NaN they make is group g's
            for con in group_constraints:
                con.apply(ws, g, g)
            observer.update_group(state, props.send_every_group)

        bad = ws.finite_group()
        if bad is not None:
            raise ValueError(f"NaN or inf encountered in object or probe, group {bad}")
        red = ws.red.cpu().numpy()
        losses = red[:, model.L['loss']]
        loss = float(numpy.mean(losses.astype(numpy.float32)))

        if want_pos:
            grads = numpy.zeros((n_pos, 2), dtype=numpy.float32)
            scale = numpy.zeros((n_pos, 1), dtype=numpy.float32)
            o = model.L['pos_num']
            for g in range(G):
                n = ws.sizes_host[g]
                grads[ws.table_host[g, :n]] = red[g, o:o + 2 * n].reshape(n, 2)
                scale[ws.table_host[g, :n]] = probe_int * n
            for s in iter_solvers:
                if 'positions' not in s.params:
                    continue
                upd = s.step_host('positions', grads, scale, loss).astype(numpy.float32)
                upd -= numpy.mean(upd, axis=0)  # apply_update (gradient/run.py:112-117)
                state.scan += upd.reshape(state.scan.shape)
            state.object.sampling.check_scan(state.scan, state.probe.sampling.extent / 2.)
            ws.set_positions(state.object.sampling, state.scan)

        # iter constraints (run.py:276-279)
        for con in iter_constraints:
            con.apply(ws)

        state.progress.iters = numpy.concatenate([state.progress.iters, [i + start_i]])
        state.progress.detector_errors = numpy.concatenate([state.progress.detector_errors, [loss]])
        observer.update_iteration(state, i, props.niter, loss)

    observer.finish_engine(state)
    return state
