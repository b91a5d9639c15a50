"""The step-chunking of the closure runs (closure._advance) on a stand-in
for a simulation (no GPU): it lands exactly on t_end while each CFL step is
at most 1.25 times the last, and it raises when a chunk passes t_end."""
import pytest

from spectral2d_amd.closure import _advance


class FakeSim:
    """A run whose CFL dt grows by `growth` per step."""

    def __init__(self, dt0, growth):
        self.t, self.n, self.dt, self.g = 0.0, 0, dt0, growth
        self.last = 0.0

    def get_time(self):
        return self.t, self.n, self.last

    def ComputeTimeStep(self, cfl):
        return self.dt

    def cfl_steps(self, k):
        for _ in range(k):
            self.t += self.dt
            self.last = self.dt
            self.dt *= self.g
            self.n += 1

    def fixed_step(self, h):
        self.t += h
        self.last = h
        self.n += 1


@pytest.mark.parametrize("growth", [1.0, 1.001, 0.9999])
def test_lands_on_t_end(growth):
    s = FakeSim(1e-3, growth)
    n = _advance(s, 0.7, 0.5, 100, 10**6, s.cfl_steps, s.fixed_step)
    assert n == s.n and abs(s.t - 0.7) <= 1e-12


def test_chunk_past_t_end_raises():
    s = FakeSim(1e-3, 1.2)  # compounding 1.2x per step: a chunk overshoots
    with pytest.raises(RuntimeError, match="passed t_end"):
        _advance(s, 5.0, 0.5, 100, 10**6, s.cfl_steps, s.fixed_step)


def test_step_limit_raises():
    s = FakeSim(1e-6, 1.0)
    with pytest.raises(RuntimeError, match="no arrival"):
        _advance(s, 1.0, 0.5, 10, 50, s.cfl_steps, s.fixed_step)
