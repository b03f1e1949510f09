"""The CPU oracle against the reference's own fixtures (tests/golden/fixtures.json, extracted from
Cuda-CMM/tests/*/data by tests/golden/extract_fixtures.py).

Tolerances follow the reference's harness, which states 1e-10..1e-12 across architectures
(Cuda-CMM/tests/test_functions.py:68-71): statistics 1e-10 relative; E_inc (1 - det, a
cancellation) 1e-11 absolute per step; warm-up errors as logged (3 significant digits).
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))
H256 = O.TWO_PI / float(np.float32(256))


@pytest.fixture(scope="module")
def one_step():
    return O.euler_run(O.grid(256), O.grid(512), O.grid(256), O.grid(512), dt=1.0, final_time=1.0,
                       skip_remapping=1, want=("chix", "chiy", "psi", "w_herm", "vort_coarse"))


def close(a, b, rel=1e-10):
    return abs(a - b) <= rel * abs(b)


def test_one_step_warmup_log(one_step):
    ref = FIX["one_step"]["log_warmup_incomp"]
    assert len(one_step["warmup_incomp"]) == len(ref)
    for a, b in zip(one_step["warmup_incomp"], ref):
        assert abs(a - b) <= 0.006 * abs(b)


def test_one_step_incompressibility(one_step):
    assert close(one_step["incomp"][0], FIX["one_step"]["Error_incompressibility"][0])


def test_one_step_map_statistics(one_step):
    fx = FIX["one_step"]
    n = 256 * 256
    for name, arr in (("ChiX", one_step["chix"]), ("ChiY", one_step["chiy"])):
        v = arr[:n]
        assert close(v.min(), fx[f"Debug_globals/Map_{name}_coarse_Min"][1])
        assert close(v.max(), fx[f"Debug_globals/Map_{name}_coarse_Max"][1])
        assert close(v.sum() * H256 * H256, fx[f"Debug_globals/Map_{name}_coarse_L2"][1])
        # Hermite arrays: the minimum over all planes sits in the cross-derivative plane, built from
        # footpoint differences over eps^2 = 1e-6, where footpoint rounding (~1e-15) becomes ~1e-9
        assert abs(arr.min() - fx[f"Debug_globals/Map_{name}_H_coarse_Min"][1]) <= 1e-8


def test_one_step_psi_and_vorticity(one_step):
    fx = FIX["one_step"]
    n = 256 * 256
    psi = one_step["psi"]
    assert close(psi[:n].min(), fx["Debug_globals/Stream_function_Psi_psi_Min"][1])
    assert close(psi[:n].max(), fx["Debug_globals/Stream_function_Psi_psi_Max"][1])
    assert close(psi[n:2 * n].max(), fx["Debug_globals/Velocity_UX_psi_Max"][1])
    assert close(psi[2 * n:3 * n].max(), fx["Debug_globals/Velocity_UY_psi_Max"][1])
    w = one_step["vort_coarse"]
    assert close(w.min(), fx["Debug_globals/Vorticity_W_coarse_Min"][1])
    assert close(0.5 * H256 * H256 * np.sum(psi[n:3 * n] ** 2), fx["Mesure/Energy"][1])
    assert close(0.5 * H256 * H256 * np.sum(w ** 2), fx["Mesure/Enstrophy"][1])


def test_initial_closed_forms():
    """4_nodes: omega = cos x + cos y + 0.6 cos 2x + 0.2 cos 3x, psi = -(cos x + cos y + 0.15 cos 2x
    + 0.2/9 cos 3x) (Laplace(psi) = omega); energy pi^2 (2 + 0.09 + 0.04/9), enstrophy 2.4 pi^2."""
    r = O.euler_run(O.grid(256), O.grid(512), O.grid(256), O.grid(512), dt=1.0, final_time=1.0,
                    skip_remapping=1, max_steps=-1, l_order=1, tin=10, want=("psi", "vort_coarse"))
    n = 256 * 256
    x = np.arange(256) * H256
    X, Y = np.meshgrid(x, x)
    psi_exact = -(np.cos(X) + np.cos(Y) + 0.15 * np.cos(2 * X) + 0.2 / 9 * np.cos(3 * X))
    assert np.max(np.abs(r["psi"][:n] - psi_exact.ravel())) < 1e-12
    E = 0.5 * H256 * H256 * np.sum(r["psi"][n:3 * n] ** 2)
    assert abs(E - np.pi ** 2 * (2 + 0.09 + 0.04 / 9)) < 1e-11
    assert abs(E - FIX["one_step"]["Mesure/Energy"][0]) < 1e-11
    Z = 0.5 * H256 * H256 * np.sum(r["vort_coarse"] ** 2)
    assert abs(Z - 2.4 * np.pi ** 2) < 1e-11


def test_remapping_11b():
    fx = FIX["remap_11b"]
    r = O.euler_run(O.grid(256), O.grid(512), O.grid(256), O.grid(512), dt=1 / 32, final_time=2.0,
                    threshold=1e-4, stack_capacity=11, want=("psi", "vort_coarse"))
    assert r["steps"] == 64
    assert np.array_equal(r["time"], np.array(fx["Time_s"]))
    assert np.array_equal(r["map_counter"], np.array(fx["Map_counter"]))
    assert np.abs(r["incomp"] - np.array(fx["Error_incompressibility"])).max() <= 1e-11
    for a, b in zip(r["warmup_incomp"], fx["log_warmup_incomp"]):
        assert abs(a - b) <= 0.006 * abs(b)
    n = 256 * 256
    E = 0.5 * H256 * H256 * np.sum(r["psi"][n:3 * n] ** 2)
    Z = 0.5 * H256 * H256 * np.sum(r["vort_coarse"] ** 2)
    assert close(E, fx["Mesure/Energy"][4])
    assert close(Z, fx["Mesure/Enstrophy"][4])
