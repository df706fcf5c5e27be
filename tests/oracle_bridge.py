"""Turns a product Problem (fivonagi_amd.Problem) into the oracle's Problem,
so a test states a workload once and runs it on both sides."""
from oracle import pyoracle as po


def to_oracle(p, form=po.FORM_LITERAL, dt_seq=None, nthreads=0, T_stop=None):
    return po.Problem(p.nx, p.ny, p.dx, p.dy, p.beta, p.delta, p.cfl_max, p.cfl_wish, p.dt0,
                      p.T0, bc=p.bc, cfl_mode=p.cfl_mode, form=form, omega=p.omega,
                      edge_amp=p.edge_amp, edge_phase=p.edge_phase, taylor=p.taylor,
                      dt_seq=dt_seq, nthreads=nthreads, moving=getattr(p, "moving", None),
                      T_stop=T_stop, hifu=getattr(p, "hifu", None))
