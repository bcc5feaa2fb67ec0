"""The simulate CLI: plain (per-function) and fused (engine) modes agree with the oracle."""
import numpy as np
import pytest

from helpers import oracle_sim, rel_inf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ARGS = ["--elements", "Fe,Mn,Ni,Co,Cu", "--temperature", "873K", "--resolution", "0.025", "--c11", "240",
        "--c12", "150", "--c44", "120", "--molar_volume", "7.1e-6", "--system_dim", "2", "--mw", "3", "--ppw", "8",
        "--direction", "0,0,1", "--initial_composition", "0.1,0.2,0.2,0.3", "--kappa_value", "5e-16",
        "--synthetic_table", "--seed", "3", "--total_timesteps", "30", "--no_save_outputs"]


def test_cli_plain_and_fused_match_oracle():
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.simulate import run
    from phasefield5d_amd.solver.config import parse_args_to_config
    plain = run(ARGS + ["--mode", "plain"])
    fused = run(ARGS + ["--mode", "fused", "--precision", "fp64"])
    f32 = run(ARGS + ["--mode", "fused", "--precision", "fp32"])
    pb = build_problem(parse_args_to_config(ARGS))
    ora = oracle_sim(pb)
    X, dt, t = pb.X_init.copy(), pb.dt0, 0.0
    for _ in range(30):
        X, _, dt, t, _ = ora.step(X, dt, t)
    assert plain["steps"] == fused["steps"] == 30
    assert rel_inf(plain["X"] - pb.X_init, X - pb.X_init) < 1e-10
    assert rel_inf(fused["X"] - pb.X_init, X - pb.X_init) < 1e-10
    assert abs(plain["time"] - t) <= 1e-12 * t and abs(fused["time"] - t) <= 1e-12 * t
    assert np.abs(f32["X"] - X).max() < 1e-6
    assert plain["traces"] and "f_grad" in plain["traces"][0]
