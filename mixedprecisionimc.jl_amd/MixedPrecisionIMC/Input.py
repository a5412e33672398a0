"""Input.readInputs -- the deck parser of src/imc_input.jl:20-129.

Same line rules as the reference: a key line must match ``^\\w+ = `` (imc_input.jl:46),
``#`` starts a comment, ``[...]`` arrays may span lines, a value holding a comma is parsed
to numbers (MESHNODES/XMESHNODES/YMESHNODES always as Float64, imc_input.jl:85-86), and a
value holding ``(`` is a tuple literal.  Tuple literals are read with ``ast.literal_eval``
(literals only) instead of Julia's ``eval(Meta.parse(...))`` (imc_input.jl:106), so a deck
cannot execute code.  The host keeps every number in FP64; ``PRECISION`` selects the
tracking precision of the GPU engine.
"""
from __future__ import annotations

import ast
import re

_KEY_RE = re.compile(r"^\w+ = ")

# deck PRECISION -> engine precision code (imc_input.jl:113-129)
PRECISIONS = {"HALF": 16, "FLOAT16": 16, "SINGLE": 32, "FLOAT32": 32, "DOUBLE": 64, "FLOAT64": 64}

NODE_KEYS = ("MESHNODES", "XMESHNODES", "YMESHNODES")


def readInputs(filename: str) -> dict:
    """Reads a deck into a dict of strings / lists (imc_input.jl:20-95)."""
    params: dict = {}
    current_key = ""
    current_value: list = []
    in_array = False
    with open(filename, "r") as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            if not in_array and _KEY_RE.match(line):
                parts = line.split("=")
                current_key = parts[0].strip()
                value = parts[1].strip()
                if current_key and not value.startswith("["):
                    params[current_key] = value
                elif value.startswith("["):
                    if not value.strip().endswith("]"):
                        in_array = True
                    current_value = [value.strip("[]")]
                    if not in_array:
                        params[current_key] = current_value
            elif in_array:
                if line.strip().endswith("]"):
                    current_value.append(line.strip("]"))
                    params[current_key] = current_value
                    in_array = False
                else:
                    current_value.append(line.strip())
    precision_set(params)
    for key, value in list(params.items()):
        if key == "PRECISION":
            continue
        joined = value if isinstance(value, str) else "".join(value)
        if "," in joined:
            params[key] = separated_array_parser(value)
    return params


def separated_array_parser(input_array) -> list:
    """imc_input.jl:97-111: tuple literals -> [tuple]; otherwise comma-separated floats."""
    joined = input_array if isinstance(input_array, str) else "".join(input_array)
    if "(" in joined:
        return [ast.literal_eval(joined)]
    return [float(tok) for tok in joined.split(",") if tok.strip() != ""]


def precision_set(params: dict) -> None:
    """imc_input.jl:113-129: maps PRECISION to the engine's precision code (16/32/64)."""
    if "PRECISION" in params and isinstance(params["PRECISION"], str):
        name = params["PRECISION"].strip().upper()
        if name in PRECISIONS:
            params["PRECISION"] = PRECISIONS[name]


def get_scalar(params: dict, key: str) -> str:
    """A scalar deck value as a string (missing keys raise KeyError, as in the reference)."""
    v = params[key]
    if isinstance(v, list):
        return str(v[0])
    return str(v)
