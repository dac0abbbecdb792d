"""Shared fixtures: the product package (libcuacd.so over HIP) and the oracle
(oracle/build/liboracle.so, a CPU restatement used only as the checker)."""
import ctypes
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "cuacd-web_amd")
ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "liboracle.so")


def load_package():
    if "cuacd_web_amd" in sys.modules:
        return sys.modules["cuacd_web_amd"]
    spec = importlib.util.spec_from_file_location("cuacd_web_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cuacd_web_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcuacd.so")


@pytest.fixture(scope="session")
def cw():
    return load_package()


@pytest.fixture(scope="session")
def oracle(cw):
    from cuacd_web_amd._ffi import Api
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle not built: run __graft_entry__.build()")
    return Api(ctypes.CDLL(ORACLE_LIB, mode=ctypes.RTLD_LOCAL), "oracle_")


@pytest.fixture(scope="session")
def product(cw):
    return cw.load_product()
