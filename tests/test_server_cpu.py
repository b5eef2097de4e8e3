"""Server argument validation that runs without a GPU."""
import ctypes
import os

import pytest

import gpu_msm_amd as gm


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(gm.LIB_PATH):
        pytest.skip("library not built")
    return gm.lib()


def test_server_rejects_bad_config_before_touching_a_device(L):
    h = ctypes.c_void_p()
    cfg = gm._ServerConfig(0, 0, 8, 1 << 30, gm.PLANNER_HEURISTIC, None)  # zero slots
    assert L.gm_server_create(ctypes.byref(h), ctypes.byref(cfg), None, None) == -1
    cfg = gm._ServerConfig(0, 2, 0, 1 << 30, gm.PLANNER_HEURISTIC, None)  # zero batch
    assert L.gm_server_create(ctypes.byref(h), ctypes.byref(cfg), None, None) == -1
    cfg = gm._ServerConfig(0, 2, 8, 1 << 30, gm.PLANNER_MLP, None)  # MLP planner without weights
    assert L.gm_server_create(ctypes.byref(h), ctypes.byref(cfg), None, None) == -1
    assert L.gm_server_submit(None, 1, 0, 4, None, 5, 299, 0, None) == -1
