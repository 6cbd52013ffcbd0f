"""What the trained network gains as the initial guess of the inner iterations.

The schedule of cgli_fu_flow_amd.reports.learned, shortened: record the
increments of consecutive channel steps (48x64x48, Re_tau 180, beta 2.0),
train the network on random crops of them with Adam, then run the held-out
steps that follow from one saved state with each initial guess: the
network's, the quadratic and linear extrapolations, and the reference's
U_ = U^n (main.cpp:378-382).  Every run stops on the same absolute tolerance,
max_v er_v <= 1e-6 E_v with E the running maxima of the recording run (the
reference's rule, main.cpp:287-304, with E shared instead of per run).

Asserted: every held-out step meets that tolerance with the network's guess
(the LU-SGS iteration stays the judge); the network's first inner residual is
below the quadratic extrapolation's in every variable; it needs fewer inner
iterations than the reference's start and no more than either extrapolation.
"""
import argparse

import numpy as np
import pytest

from cgli_fu_flow_amd import reports

pytestmark = pytest.mark.gpu


def test_trained_network_guess_beats_the_guesses_without_it():
    a = argparse.Namespace(nx=48, ny=64, nz=48, re_tau=180.0, beta=2.0, skew=0.0, cfl=0.5, tol=1e-6,
                           max_inner=100, precision="fp64", gen=300, train=400, batch=8, crop=32, held=40,
                           held_tols="1e-6", held_rule="running-max",
                           guesses="network,quadratic,extrapolate,previous", seed=20261203)
    print("learned-guess schedule:", a, flush=True)
    out = reports.learned(a)
    res = out["held_out"]["1e-06"]
    runs = res["runs"]
    for g, r in runs.items():
        print("%-12s %.2f inner/step  first residual / E %s" % (g, r["inner_iterations_per_step"],
                                                                np.array2string(np.array(r["first_residual_over_E"]),
                                                                                precision=3)), flush=True)
    print("train loss %.3e -> %.3e" % (out["train"]["loss_first20"], out["train"]["loss_last20"]))
    net = runs["network"]
    assert net["steps_meeting_tolerance"] == a.held
    assert net["network_guess_taken"] > 0
    assert max(res["network_first_residual_vs_quadratic"]) < 1.0
    assert net["inner_iterations_per_step"] < runs["previous"]["inner_iterations_per_step"]
    assert net["inner_iterations_per_step"] <= runs["extrapolate"]["inner_iterations_per_step"]
    assert net["inner_iterations_per_step"] <= runs["quadratic"]["inner_iterations_per_step"]
