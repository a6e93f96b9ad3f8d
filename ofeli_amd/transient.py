"""Backward-Euler time stepping of DC2DT3 with device-resident state.

The reference drives a transient heat problem through TimeStepping with
BACKWARD_EULER (TimeStepping.cpp:174-212, 445-456) over DC2DT3, whose
per-step system is built by Equa_Therm::build (Equa_Therm.h:162-218):
    (C/dt + K) u^{n+1} = C/dt u^n + F^{n+1}
with C the lumped capacity (DC2DT3::LCapacity) and K the diffusion matrix.
Here the matrix is assembled once on the GPU (k and dt are fixed), each step
assembles only the right-hand side on the GPU from the nodal u^n, and the
solution never leaves the device.  The start vector of each solve is the
previous solution (the reference's warm start, TimeStepping.cpp:207), a
linear extrapolation, or a LearnedGuess.
"""
from . import api
from . import DC2DT3_BE, DIAG_PREC


class HeatBE:
    """judge_every > 0 lets plain projection judge the network: every
    judge_every-th step with a (ProjectionGuess, LearnedGuess) start computes
    both starts and keeps the network for the following steps only if its
    start's residual is below judge_keep times the projection's; otherwise
    the following steps use the projection alone (the network's inference
    is skipped until the next judged step)."""

    def __init__(self, ctx, coord, elem, code, k, dt, rho=1.0, cp=1.0, judge_every=0,
                 judge_keep=0.5):
        import torch
        self.P = api.AssemblyPlan(ctx, coord, elem, code)
        self.A = self.P.A
        self.n = self.A.n
        self.k = torch.as_tensor(k, dtype=torch.float64, device="cuda")
        self.prm = (float(dt), float(rho), float(cp))
        z = lambda m: torch.zeros(m, dtype=torch.float64, device="cuda")
        self.u = z(coord.shape[0])   # nodal u^n
        self.x = z(self.n)           # free DOFs of u^n
        self.xp = z(self.n)          # free DOFs of u^{n-1}
        self.b = z(self.n)
        self.x0 = z(self.n)          # start vector, then the new solution
        self.P.build(DC2DT3_BE, coef=self.k, f=z(coord.shape[0]), u_prev=self.u, prm=self.prm,
                     matrix=True, b=self.b)
        self.step_count = 0
        self.judge_every, self.judge_keep = int(judge_every), float(judge_keep)
        self.use_net = True
        self.judged = 0      # judged steps
        self.net_steps = 0   # steps started from the network's guess

    def rhs(self, f):
        """b = C/dt u^n + F(f) (Equa_Therm.h:197-205), on the GPU."""
        self.P.build(DC2DT3_BE, coef=self.k, f=f, u_prev=self.u, prm=self.prm, matrix=False,
                     b=self.b)

    def start(self, guess=None):
        """Start vector of the next solve in x0: "warm" (u^n), "extrapolate"
        (2 u^n - u^{n-1}), a ProjectionGuess (span of previous solutions), a
        LearnedGuess, or (ProjectionGuess, LearnedGuess): the network
        corrects the projection (its base x_b is the projected vector, its
        "previous" iterate u^n)."""
        if guess is None or guess == "warm":
            self.x0.copy_(self.x)
        elif guess == "extrapolate":
            self.x0.copy_(self.x).mul_(2.0).sub_(self.xp)
        elif isinstance(guess, api.ProjectionGuess):
            guess.guess(self.b, self.x0)
        elif isinstance(guess, tuple):
            proj, net = guess
            if not hasattr(self, "xg"):
                self.xg = self.x0.clone()
                self.rq = self.x0.clone()
            proj.guess(self.b, self.xg)
            judge = self.judge_every > 0 and self.step_count % self.judge_every == 0
            if judge:
                net.guess(self.xg, self.x, self.b, self.x0)
                rn, rp = self._resnorm(self.x0), self._resnorm(self.xg)
                self.use_net = rn < self.judge_keep * rp
                self.judged += 1
                if not self.use_net:
                    self.x0.copy_(self.xg)
            elif self.use_net or self.judge_every <= 0:
                net.guess(self.xg, self.x, self.b, self.x0)
            else:
                self.x0.copy_(self.xg)
            if self.use_net or self.judge_every <= 0:
                self.net_steps += 1
        else:
            guess.guess(self.x, self.xp, self.b, self.x0)

    def _resnorm(self, x):
        """||b - A x||, on the GPU."""
        import torch
        api.spmv(self.A, x, self.rq)
        return float(torch.linalg.norm(self.b - self.rq))

    def solve(self, mixed=False, toler=1e-8, max_it=20000, prec=DIAG_PREC):
        fn = api.cg_mixed if mixed else api.cg
        return fn(self.A, self.b, self.x0, prec=prec, max_it=max_it, toler=toler)

    def advance(self, guess=None):
        self.xp.copy_(self.x)
        self.x.copy_(self.x0)
        self.P.insert_bc(self.x, None, self.u)
        self.step_count += 1
        proj = guess[0] if isinstance(guess, tuple) else guess
        if isinstance(proj, api.ProjectionGuess):
            proj.update(self.x)

    def step(self, f, guess=None, mixed=False, toler=1e-8):
        self.rhs(f)
        self.start(guess)
        info = self.solve(mixed=mixed, toler=toler)
        self.advance(guess)
        return info

    def state(self):
        return (self.u.clone(), self.x.clone(), self.xp.clone(), self.step_count)

    def set_state(self, st):
        u, x, xp, c = st
        self.u.copy_(u)
        self.x.copy_(x)
        self.xp.copy_(xp)
        self.step_count = c
