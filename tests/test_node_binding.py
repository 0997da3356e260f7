"""The N-API binding of libmdcore.so under Node (the host upstream runs in).

The addon (``node/mdcore_node.c``) is compiled with the system C compiler
against Node's own headers, and ``node/check.js`` drives it through
``node/engine.js``, the ``ProtocolEngine`` a TypeScript host uses:

* CPU lane: it loads, reports its version, returns the cell sort order, and
  raises upstream's named errors (CutoffTooLargeForBox, NotSmooth,
  InvalidArgument) from ``createSystem``; without a GPU a valid System fails
  as NoDevice -- no fallback.
* GPU lane: the 10-particle OpenMM fixture's direct and reciprocal group
  energies within 1e-5 of OpenMM's, ``run`` as a Promise that advances the
  step counter, a sample with finite energies, and a named refusal from
  ``setBox`` after which the engine keeps running.
"""
import json
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT
from mdcore import native
from mdcore.state import parseState
from mdcore.system import parseSystem

NODE_DIR = os.path.join(ROOT, "making-it-rain-webgpu_amd", "node")
ADDON = os.path.join(NODE_DIR, "mdcore.node")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists("/usr/include/node/node_api.h"),
                                reason="Node or its headers are not installed")


def _fixture(tmp_path):
    system = parseSystem(open(os.path.join(GOLDEN, "nonbonded-system.xml")).read())
    state = parseState(open(os.path.join(GOLDEN, "nonbonded-state.xml")).read())
    facts = json.load(open(os.path.join(GOLDEN, "nonbonded-facts.json")))
    desc = {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in native.describe_fields(system).items()}
    desc["grid"] = list(facts["PME"]["grid"])
    path = tmp_path / "fixture.json"
    path.write_text(json.dumps({"desc": desc, "positions": state.positions.tolist(),
                                "facts": {"direct_group": facts["PME"]["direct_group"],
                                          "reciprocal_group": facts["PME"]["reciprocal_group"]}}))
    return str(path)


def _build():
    subprocess.run(["make", "-s", "-C", os.path.dirname(NODE_DIR), "node"], check=True)
    assert os.path.exists(ADDON)


def _check(mode, fixture):
    r = subprocess.run([NODE, os.path.join(NODE_DIR, "check.js"), mode, fixture], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert f"node binding ({mode}): ok" in r.stdout


def test_binding_loads_and_raises_named_errors_without_a_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; the GPU lane covers the binding")
    _build()
    _check("cpu", _fixture(tmp_path))


@pytest.mark.gpu
def test_binding_runs_the_fixture_on_the_gpu(tmp_path):
    assert os.path.exists(ADDON), "build the addon first (make -C making-it-rain-webgpu_amd node)"
    _check("gpu", _fixture(tmp_path))
