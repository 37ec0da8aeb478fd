"""Pipeline-schema overlay for the library's nodes, generated from their manifests.

The reference's pipeline schema constrains each node's params with an
`allOf` list of `if type == <node id> then params {...}` blocks; for
com.cpipe.denoise.bm3d it allows `sigma` (0..0.2) and `sigma_override` (>= 0)
only (schemas/pipeline-v0.4.json:101-117), so a pipeline file cannot select the
`mode`, `precision`, `pilot` or `noise_map` this node adds. This module derives the
successor blocks for the three nodes of the library from the manifests the
library embeds (cpipe_amd/csrc/*.json), adds the rules a manifest cannot state
(pilot=learned needs mode=bm3d; noise_map=profile needs mode=bm3d and no pilot;
precision_convert's target_format, which the
reference's precision planner sets, precision_convert.cpp:54-57), and writes
them to cpipe_amd/schemas/pipeline-v0.5-nodes.json, a fragment a host merges
into its pipeline schema's node `allOf`. The build regenerates it.

`validate` checks an instance against the keyword subset the overlay uses
(type, enum, const, minimum, maximum, properties, required,
additionalProperties, allOf, if/then), since no JSON-Schema library is
installed here.
"""
from __future__ import annotations

import json
import os

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "schemas", "pipeline-v0.5-nodes.json")
MANIFESTS = ("bm3d.json", "wavelet_bayes_shrink.json", "precision_convert.json")

# Params that a host, not the pipeline author, sets (manifests do not list them).
HOST_SET = {"com.cpipe.precision_convert": {
    "target_format": {"enum": ["R32_SFLOAT", "R16G16B16A16_SFLOAT", "R8G8B8A8_UNORM", "R16G16B16A16_UNORM"]}}}
# Cross-parameter rules the node enforces (CPIPE_BAD_INDEX otherwise).
RULES = {"com.cpipe.denoise.bm3d": [
    {"if": {"properties": {"pilot": {"const": "learned"}}, "required": ["pilot"]},
     "then": {"properties": {"mode": {"const": "bm3d"}}, "required": ["mode"]}},
    {"if": {"properties": {"noise_map": {"const": "profile"}}, "required": ["noise_map"]},
     "then": {"properties": {"mode": {"const": "bm3d"}, "pilot": {"const": "off"}}, "required": ["mode"]}}]}


def _param_schema(p: dict) -> dict:
    if p["type"] == "number":
        s = {"type": "number"}
        rng = p.get("range", {})
        if "min" in rng:
            s["minimum"] = rng["min"]
        if "max" in rng:
            s["maximum"] = rng["max"]
        return s
    if p["type"] == "enum":
        return {"enum": list(p["enum_values"])}
    raise ValueError(f"param type {p['type']!r} has no schema mapping")


def node_block(manifest: dict) -> dict:
    props = {p["name"]: _param_schema(p) for p in manifest.get("params", [])}
    props.update(HOST_SET.get(manifest["id"], {}))
    params = {"type": "object", "additionalProperties": False, "properties": props}
    rules = RULES.get(manifest["id"])
    if rules:
        params["allOf"] = rules
    return {"if": {"properties": {"type": {"const": manifest["id"]}}, "required": ["type"]},
            "then": {"properties": {"params": params}}}


def overlay() -> dict:
    blocks = []
    for name in MANIFESTS:
        with open(os.path.join(CSRC, name), encoding="utf-8") as f:
            blocks.append(node_block(json.load(f)))
    return {
        "$schema": "https://json-schema.org/draft/2020-12/schema",
        "$id": "cpipe_amd/schemas/pipeline-v0.5-nodes.json",
        "title": "Node params of com.cpipe.denoise.bm3d, com.cpipe.denoise.wavelet_bayes_shrink and "
                 "com.cpipe.precision_convert (pipeline schema 0.5 node blocks)",
        "description": "Generated from the nodes' manifests by cpipe_amd/schema.py; merge into the "
                       "pipeline schema's node items allOf. Successor of the bm3d block of pipeline "
                       "schema 0.4, adding mode, precision, pilot and noise_map.",
        "allOf": blocks,
    }


def render() -> str:
    return json.dumps(overlay(), indent=2) + "\n"


def write(path: str = OUT) -> str:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    text = render()
    old = None
    if os.path.exists(path):
        with open(path, encoding="utf-8") as f:
            old = f.read()
    if old != text:
        with open(path, "w", encoding="utf-8") as f:
            f.write(text)
    return path


_TYPES = {"object": dict, "string": str, "array": list, "boolean": bool}


def _type_ok(value, t: str) -> bool:
    if t == "number":
        return isinstance(value, (int, float)) and not isinstance(value, bool)
    if t == "integer":
        return isinstance(value, int) and not isinstance(value, bool)
    return isinstance(value, _TYPES[t])


def validate(instance, schema: dict, path: str = "$") -> list[str]:
    """Errors of `instance` against `schema` (empty when valid)."""
    errs: list[str] = []
    if "type" in schema and not _type_ok(instance, schema["type"]):
        return [f"{path}: expected {schema['type']}"]
    if "const" in schema and instance != schema["const"]:
        errs.append(f"{path}: must be {schema['const']!r}")
    if "enum" in schema and instance not in schema["enum"]:
        errs.append(f"{path}: {instance!r} not in {schema['enum']}")
    if isinstance(instance, (int, float)) and not isinstance(instance, bool):
        if "minimum" in schema and instance < schema["minimum"]:
            errs.append(f"{path}: {instance} < {schema['minimum']}")
        if "maximum" in schema and instance > schema["maximum"]:
            errs.append(f"{path}: {instance} > {schema['maximum']}")
    if isinstance(instance, dict):
        for key in schema.get("required", []):
            if key not in instance:
                errs.append(f"{path}: missing {key!r}")
        props = schema.get("properties", {})
        for key, value in instance.items():
            if key in props:
                errs += validate(value, props[key], f"{path}.{key}")
            elif schema.get("additionalProperties", True) is False:
                errs.append(f"{path}: unknown property {key!r}")
    for sub in schema.get("allOf", []):
        errs += validate(instance, sub, path)
    if "if" in schema and not validate(instance, schema["if"], path) and "then" in schema:
        errs += validate(instance, schema["then"], path)
    return errs


if __name__ == "__main__":
    print(write())
