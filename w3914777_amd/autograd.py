"""torch.autograd binding of the GPU adjoint (thin: used only for gradients).

``PhaseFieldRun.apply(params, cfg, param_ids, nsteps, ckpt_every, init)``
runs a PURE_FP64 context of ``cfg`` for ``nsteps`` explicit-Euler steps with
the parameters named by ``param_ids`` (``GRAD_SF`` / ``GRAD_SN`` ids) set to
the values of the 1-D tensor ``params``, recording a checkpoint every
``ckpt_every`` steps, and returns the final fields as a float64 CPU tensor
``[nfields, ny, nx]``.  ``init`` (optional ``[nfields, ny, nx]`` tensor)
replaces the model's initial fields.  ``backward`` runs the reverse sweep of
the recorded run on the GPU (``amps_adjoint``) and returns d(loss)/d(params)
and, when ``init`` requires grad, d(loss)/d(init).

No reference counterpart (the reference has no differentiation).
"""
import numpy as np
import torch

from . import GRAD_SF, GRAD_SN, PURE_FP64, SF, SN, Context

# gradient id -> parameter slot of amps_config.params
PARAM_SLOT = {
    GRAD_SF["DELTA"]: SF["DELTA"], GRAD_SF["KAPPA"]: SF["KAPPA"], GRAD_SF["EPSB"]: SF["EPSB"],
    GRAD_SF["TAU"]: SF["TAU"],
    GRAD_SN["DSURF"]: SN["DSURF"], GRAD_SN["DGB"]: SN["DGB"], GRAD_SN["KC"]: SN["KC"], GRAD_SN["KE"]: SN["KE"],
    GRAD_SN["L"]: SN["L"], GRAD_SN["DVOL"]: SN["DVOL"],
}


def config_with(cfg, param_ids, values):
    """Copy of cfg (PURE_FP64) with the parameters of param_ids set to values."""
    c = cfg.copy()
    c.mode = PURE_FP64
    for pid, v in zip(param_ids, values):
        if pid not in PARAM_SLOT:
            raise ValueError(f"parameter id {pid} is not differentiable")
        c.params[PARAM_SLOT[pid]] = float(v)
    return c


class PhaseFieldRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, cfg, param_ids, nsteps, ckpt_every, init=None):
        ids = [int(i) for i in param_ids]
        vals = params.detach().to("cpu", torch.float64).reshape(-1).tolist()
        if len(vals) != len(ids):
            raise ValueError("params and param_ids differ in length")
        run = Context(config_with(cfg, ids, vals))
        try:
            run.init()
            if init is not None:
                arr = init.detach().to("cpu", torch.float64).numpy()
                for f in range(arr.shape[0]):
                    run.set_field(f, arr[f])
            run.tape_begin(int(ckpt_every))
            run.step(int(nsteps))
            run.sync()
            out = np.stack([run.get_field(f) for f in range(cfg.nfields())])
        except Exception:
            run.close()
            raise
        ctx.run = run
        ctx.ids = ids
        ctx.want_init = init is not None and init.requires_grad
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, grad_out):
        run = ctx.run
        try:
            g = grad_out.detach().to("cpu", torch.float64).contiguous().numpy()
            dp, du0 = run.adjoint([g], ctx.ids, want_du0=ctx.want_init)
        finally:
            run.close()
            ctx.run = None
        gp = torch.from_numpy(np.ascontiguousarray(dp[0]))
        gi = torch.from_numpy(np.ascontiguousarray(du0[0])) if ctx.want_init else None
        return gp, None, None, None, None, gi


def run_fields(params, cfg, param_ids, nsteps, ckpt_every=100, init=None):
    """Differentiable final fields of a PURE_FP64 run (see PhaseFieldRun)."""
    return PhaseFieldRun.apply(params, cfg, param_ids, nsteps, ckpt_every, init)
