import os
import sys
import threading

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLD = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def quad_rule(ng, ngf):
    q = gold("quadrature.npz")
    t, g = q[f"tri{ng}"], q[f"legen{ngf}"]
    return t[0], t[1], t[2], g[0], g[1]


def host_tables(nc, ng, ngf):
    """Operator tables from the product's host builder (no GPU needed)."""
    import spoc_amd
    return spoc_amd.build_tables(nc, *quad_rule(ng, ngf))


def oracle_for(nc, ng, ngf):
    import oracle
    O = oracle.Oracle()
    O.setup(nc, *quad_rule(ng, ngf))
    return O


def big_stack(fn, *a, **k):
    """Run fn on a thread with a large stack (the reference's nonlinear
    variants keep big work arrays on the stack)."""
    res = {}

    def tgt():
        try:
            res["v"] = fn(*a, **k)
        except BaseException as e:  # re-raised below
            res["e"] = e

    old = threading.stack_size(512 * 1024 * 1024)
    try:
        t = threading.Thread(target=tgt)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in res:
        raise res["e"]
    return res["v"]


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def have_gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    return True
