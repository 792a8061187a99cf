"""Gradients through many MPDATA steps: a torch.autograd.Function over the
hand-written adjoint kernels of libmpdata_hip (non-oscillatory scheme,
standard or gauge variant, 2 iterations, open walls, FP64).

Forward runs ``nsteps`` steps and keeps the field every ``ckpt`` steps; the
steps of the last segment run as recorded steps, which keep their
intermediates on the library's tape (one slot per step of a segment).
Backward walks the checkpoint segments from the last: the last one's tape is
already there (unless another recording on the domain replaced it); every
other segment is recomputed from its checkpoint with recording, on a second
stream while the adjoint steps of the segment after it run, into the other
half of a tape of two segments.  Memory: ``nsteps/ckpt`` fields plus
``2 ckpt`` tape slots.

The velocity of step n is ``factors[n] * (u1, u2, u3)`` (all ones by
default), so a time-dependent flow of fixed shape, such as the reversing
deformational flow, has gradients with respect to its shape.
"""
from __future__ import annotations

import torch

from . import mpdata


def _scaled(u, f):
    return u if f == 1.0 else u * f


def _complete_halo(x, x_prev, x0):
    """The halo of x (ih = 1) as a plain step from x_prev leaves it, after a
    recorded step, which defines only the interior: edge copies of x_prev (the
    field at the step's time level n) on the x/y walls of the interior planes,
    the other halo cells (lid planes, x/y corners) those of the input x0,
    which no step changes."""
    x[0].copy_(x0[0])
    x[-1].copy_(x0[-1])
    m = slice(1, -1)
    x[m, m, 0].copy_(x_prev[m, m, 1])
    x[m, m, -1].copy_(x_prev[m, m, -2])
    x[m, 0, m].copy_(x_prev[m, 1, m])
    x[m, -1, m].copy_(x_prev[m, -2, m])
    for jj in (0, -1):
        for ii in (0, -1):
            x[m, jj, ii].copy_(x0[m, jj, ii])


class MpdataSteps(torch.autograd.Function):
    """x_T = MPDATA^nsteps(x_0; u).  Differentiable in x0, u1, u2, u3 (the
    interiors; halo cells of x carry no gradient)."""

    @staticmethod
    def forward(ctx, x0, u1, u2, u3, dom, rhr, h, bcx, bcy, nsteps, ckpt, factors, variant):
        factors = [1.0] * nsteps if factors is None else [float(f) for f in factors]
        if len(factors) != nsteps:
            raise ValueError("one velocity factor per step")
        ckpts = []
        x = x0.detach().clone()
        # the last segment's steps are recorded, so that the backward starts
        # with its tape
        taped = any(ctx.needs_input_grad[:4])
        last0 = ((nsteps - 1) // ckpt) * ckpt if nsteps > 0 else 0
        base = _tape_base(last0 // ckpt, ckpt)
        if taped and nsteps > 0:
            dom.adj_reserve(_tape_slots(nsteps, ckpt))
        spare = None
        last_vel = []
        for n in range(nsteps):
            if n % ckpt == 0:
                ckpts.append(x.clone())
            un = [_scaled(u, factors[n]) for u in (u1, u2, u3)]
            if taped and n >= last0:
                if spare is None:
                    spare = torch.empty_like(x)
                dom.adj_record(base + n - last0, x, spare, *un, rhr, h, bcx, bcy, variant=variant)
                x, spare = spare, x
                last_vel.append(un)
            else:
                dom.step(x, *un, rhr, h, bcx, bcy, variant=variant, iters=2, nonos=1)
        if taped and nsteps > 0:
            _complete_halo(x, spare, x0.detach())
        ctx.tape_epoch = dom.tape_epoch if taped else None
        ctx.last_vel = last_vel
        ctx.dom, ctx.nsteps, ctx.ckpt, ctx.factors = dom, nsteps, ckpt, factors
        ctx.variant = variant
        ctx.ckpts = ckpts
        ctx.fields = (u1.detach(), u2.detach(), u3.detach(), rhr, h, bcx, bcy)
        return x

    @staticmethod
    def backward(ctx, grad):
        dom, nsteps, ckpt, factors = ctx.dom, ctx.nsteps, ctx.ckpt, ctx.factors
        u1, u2, u3, rhr, h, bcx, bcy = ctx.fields
        dom.adj_reserve(_tape_slots(nsteps, ckpt))
        main = torch.cuda.current_stream(grad.device)
        side = torch.cuda.Stream(grad.device)  # default priority
        xbar = torch.zeros_like(grad)
        mpdata.interior(xbar).copy_(mpdata.interior(grad))
        ub = [torch.zeros_like(u) for u in (u1, u2, u3)]
        bufs = [torch.zeros_like(grad), torch.zeros_like(grad)]
        nseg = len(ctx.ckpts)
        # the inputs the side stream reads were written on the main stream
        ready = torch.cuda.Event()
        ready.record(main)
        side.wait_event(ready)

        def record(seg):
            """Recomputes segment seg with recording (side stream), into the
            tape half of its parity; returns its velocities."""
            s0, s1 = seg * ckpt, min(nsteps, seg * ckpt + ckpt)
            base = _tape_base(seg, ckpt)
            with torch.cuda.stream(side):
                vel = [[_scaled(u, factors[n]) for u in (u1, u2, u3)] for n in range(s0, s1)]
                x = ctx.ckpts[seg]
                for n in range(s0, s1):
                    out = bufs[(n - s0) & 1]
                    dom.adj_record(base + n - s0, x, out, *vel[n - s0], rhr, h, bcx, bcy,
                                   variant=ctx.variant)
                    x = out
            for v in vel:
                for t in v:
                    if t.data_ptr() not in (u1.data_ptr(), u2.data_ptr(), u3.data_ptr()):
                        t.record_stream(main)  # read by the adjoint steps on main
            done = torch.cuda.Event()
            done.record(side)
            return vel, done

        # the forward recorded the last segment (and kept its velocities); its
        # tape is intact if nothing has recorded on this domain since
        pending = None
        if ctx.tape_epoch is not None and dom.tape_epoch == ctx.tape_epoch:
            cur = (ctx.last_vel, None)
        else:
            cur = record(nseg - 1)
        freed = None  # main-stream event: the tape half of parity seg+1 is free
        for seg in reversed(range(nseg)):
            s0, s1 = seg * ckpt, min(nsteps, seg * ckpt + ckpt)
            if seg >= 1:
                # segment seg-1 overwrites the tape half of segment seg+1
                if freed is not None:
                    side.wait_event(freed)
                pending = record(seg - 1)
            vel, done = cur
            if done is not None:
                main.wait_event(done)
            base = _tape_base(seg, ckpt)
            for n in reversed(range(s0, s1)):
                # u_n = f_n u: the gradient of u collects f_n times that of u_n
                dom.adj_step(base + n - s0, *vel[n - s0], rhr, h, xbar, *ub, ufac=factors[n])
            freed = torch.cuda.Event()
            freed.record(main)
            cur = pending
        main.wait_stream(side)
        ctx.ckpts = None
        ctx.last_vel = None
        return (xbar, ub[0], ub[1], ub[2]) + (None,) * 9


def _tape_base(seg, ckpt):
    """First tape slot of checkpoint segment seg: consecutive segments use
    the two halves of the tape alternately."""
    return (seg % 2) * ckpt


def _tape_slots(nsteps, ckpt):
    nseg = (nsteps + ckpt - 1) // ckpt
    return min(ckpt, nsteps) if nseg <= 1 else 2 * ckpt


def advect(x0, u1, u2, u3, dom, rhr, h, bcx, bcy, nsteps, ckpt=10, factors=None,
           variant=mpdata.STANDARD):
    """nsteps non-oscillatory MPDATA steps (standard or gauge variant, 2
    iterations), differentiable in x0 and the face velocities, with
    checkpoints every `ckpt` steps."""
    if variant not in (mpdata.STANDARD, mpdata.GAUGE):
        raise ValueError("the adjoint covers the standard and gauge variants")
    return MpdataSteps.apply(x0, u1, u2, u3, dom, rhr, h, bcx, bcy, nsteps, ckpt, factors,
                             variant)
