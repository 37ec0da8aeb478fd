"""The pipeline-schema node blocks generated from the manifests (cpipe_amd/schema.py), no GPU.

Checks: the committed overlay is what the manifests generate; the reference's
own pipeline node entries (schema 0.4) stay valid; entries using the added
params are valid and out-of-range or unknown ones are not; the reference's
0.4 bm3d block (schemas/pipeline-v0.4.json:101-117: sigma in [0, 0.2],
sigma_override >= 0, nothing else) is kept as a subset; and the node's own
parameter parsing accepts exactly the enum values the schema lists.
"""
import json
import os

import pytest

from cpipe_amd import host, schema

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _entries(kind):
    with open(os.path.join(GOLDEN, "pipeline_nodes.json")) as f:
        return json.load(f)[kind]


def _blocks():
    return {b["if"]["properties"]["type"]["const"]: b["then"]["properties"]["params"]
            for b in schema.overlay()["allOf"]}


def test_committed_overlay_is_generated():
    with open(schema.OUT) as f:
        assert f.read() == schema.render()


@pytest.mark.parametrize("kind", ["schema_0_4", "schema_0_5"])
def test_pipeline_node_entries_validate(kind):
    s = schema.overlay()
    for node in _entries(kind):
        assert schema.validate(node, s) == [], node


@pytest.mark.parametrize("node_type,params", [
    ("com.cpipe.denoise.bm3d", {"precision": "int4"}),
    ("com.cpipe.denoise.bm3d", {"sigma": 0.3}),
    ("com.cpipe.denoise.bm3d", {"sigma_override": -0.1}),
    ("com.cpipe.denoise.bm3d", {"mode": "fast"}),
    ("com.cpipe.denoise.bm3d", {"pilot": "learned"}),
    ("com.cpipe.denoise.bm3d", {"pilot": "learned", "mode": "compat"}),
    ("com.cpipe.denoise.bm3d", {"noise_map": "profile"}),
    ("com.cpipe.denoise.bm3d", {"noise_map": "profile", "mode": "bm3d", "pilot": "learned"}),
    ("com.cpipe.denoise.bm3d", {"noise_map": "on", "mode": "bm3d"}),
    ("com.cpipe.denoise.bm3d", {"strength": 1.0}),
    ("com.cpipe.denoise.bm3d", {"sigma": "0.05"}),
    ("com.cpipe.denoise.wavelet_bayes_shrink", {"chroma_strength": 2.5}),
    ("com.cpipe.precision_convert", {"target_format": "R8_UNORM"}),
])
def test_bad_params_are_rejected(node_type, params):
    errs = schema.validate({"id": "n", "type": node_type, "params": params}, schema.overlay())
    assert errs, params


def test_reference_bm3d_block_is_a_subset():
    bm3d = _blocks()["com.cpipe.denoise.bm3d"]["properties"]
    assert bm3d["sigma"] == {"type": "number", "minimum": 0, "maximum": 0.2}
    assert bm3d["sigma_override"] == {"type": "number", "minimum": 0}
    assert set(bm3d) == {"sigma", "sigma_override", "mode", "precision", "pilot", "noise_map"}


def test_node_parses_exactly_the_schema_enums():
    """Without a device the node parses its params, then answers
    CPIPE_UNSUPPORTED; a value it rejects gives CPIPE_BAD_INDEX first."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("this checks the no-device status")
    props = _blocks()["com.cpipe.denoise.bm3d"]["properties"]
    d = host.find_node()

    def status(params):
        a = host.Buffer(host.FORMAT_RGBA16F, 4, 4)
        a.set_noise_profile([(1e-4, 0.01)] * 3)
        b = host.Buffer(host.FORMAT_RGBA16F, 4, 4)
        return host.run_node(d, [a], [b], params)

    s = schema.overlay()
    for mode in props["mode"]["enum"]:
        for prec in props["precision"]["enum"]:
            for pilot in props["pilot"]["enum"]:
                for nmap in props["noise_map"]["enum"]:
                    params = {"mode": mode, "precision": prec, "pilot": pilot, "noise_map": nmap}
                    valid = not schema.validate({"type": "com.cpipe.denoise.bm3d", "params": params}, s)
                    want = host.CPIPE_UNSUPPORTED if valid else host.CPIPE_BAD_INDEX
                    assert status(params) == want, params
    for key, bad in (("mode", "fast"), ("precision", "int4"), ("pilot", "on"), ("noise_map", "on")):
        assert status({key: bad}) == host.CPIPE_BAD_INDEX
