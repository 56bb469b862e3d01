"""The C ABI library: it is built for gfx950 and exports every function include/meteolab.h declares (no GPU needed:
nothing here calls into the device)."""
import os
import re

import pytest

import meteolab_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "meteolab.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ml_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_bound_signatures():
    assert declared() == sorted(M.SIGNATURES)


def test_library_loads_and_exports_every_declared_symbol():
    lib = M.lib()
    for name in declared():
        assert hasattr(lib, name), name


def test_library_holds_gfx950_code():
    # The fat binary names its code objects by target triple and processor.
    data = open(M.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_field_ids_match_the_header():
    text = open(os.path.join(ROOT, "include", "meteolab.h")).read()
    body = text[text.index("ML_RAIN = 0"):text.index("ML_FIELDS")]
    names = re.findall(r"ML_([A-Z_]+)", body)
    names = [n for n in names if n not in ("CARRIED",)]
    seen = []
    for n in names:
        if n not in seen:
            seen.append(n)
    camel = [re.sub(r"_([a-z])", lambda m: m.group(1).upper(), n.lower()) for n in seen]
    fix = {"fallChargeIce": "fallChargeIce", "icenumber": "iceNumber", "fieldz": "fieldZ"}
    camel = [fix.get(c, c) for c in camel]
    assert camel == M.FIELDS
