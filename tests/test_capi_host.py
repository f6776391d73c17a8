"""The C-ABI library loads, exports every declared symbol, and its host-only
logic (profile selection) matches the reference.  No device work here."""
import ctypes as C
import re
from pathlib import Path

import pytest

from scrs_amd import _lib
from scrs_amd.codec import ConfigError, Engine, Field, Profile, recommended_gf16_engine

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols() -> set[str]:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(scrs_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    declared = declared_symbols()
    assert len(declared) > 20
    for name in declared:
        assert hasattr(lib, name), name
    assert declared <= set(_lib.SIGNATURES), sorted(declared - set(_lib.SIGNATURES))


def test_abi_version():
    assert _lib.load().scrs_abi_version() >= 1


def test_resolve_check_order():
    # selector.rs:40-61
    with pytest.raises(ConfigError) as e:
        Profile.resolve(Engine.TOWER, 0, 0, 0)
    assert e.value.variant == "ZeroDimension"
    with pytest.raises(ConfigError) as e:
        Profile.resolve(Engine.TOWER, 1, 1, 0)
    assert e.value.variant == "ZeroSymbolLen"
    with pytest.raises(ConfigError) as e:
        Profile.resolve(Engine.TOWER, 70000, 1, 3)
    assert e.value.variant == "OddSymbolLen"
    with pytest.raises(ConfigError) as e:
        Profile.resolve(Engine.TOWER, 70000, 1, 4)
    assert e.value == ConfigError("TooManySymbols", cap=65535)
    assert Profile.resolve(Engine.GOOD_CAUCHY, 3, 2, 5) == Profile(Engine.GOOD_CAUCHY, 3, 2, 5)


@pytest.mark.parametrize("engine,cap", [(Engine.STANDARD_CAUCHY, 256), (Engine.GOOD_CAUCHY, 255),
                                        (Engine.TOWER, 65535), (Engine.AFFT, 65536)])
def test_capacities_are_exact(engine, cap):
    Profile.resolve(engine, cap - 1, 1, 2)
    with pytest.raises(ConfigError) as e:
        Profile.resolve(engine, cap, 1, 2)
    assert e.value.fields == {"cap": cap}


def test_gf256_odd_symbol_len_is_fine():
    assert Profile.resolve(Engine.STANDARD_CAUCHY, 10, 4, 4097).symbol_len == 4097


def test_recommended():
    # selector.rs:67-87
    assert Profile.recommended(Field.GF256, 200, 55, 8).engine == Engine.GOOD_CAUCHY
    assert Profile.recommended(Field.GF256, 200, 56, 8).engine == Engine.STANDARD_CAUCHY
    assert Profile.recommended(Field.GF65536, 32, 8, 8).engine == Engine.TOWER
    assert Profile.recommended(Field.GF65536, 9, 3, 8).engine == Engine.AFFT
    with pytest.raises(ConfigError):
        Profile.recommended(Field.GF65536, 9, 3, 7)


def test_recommended_gf16_engine_thresholds():
    # lib.rs:95-131
    assert recommended_gf16_engine(32, 8) == Engine.TOWER
    assert recommended_gf16_engine(200, 40) == Engine.TOWER
    assert recommended_gf16_engine(4, 1) == Engine.TOWER
    assert recommended_gf16_engine(256, 1) == Engine.AFFT
    assert recommended_gf16_engine(1024, 8) == Engine.AFFT
    assert recommended_gf16_engine(9, 3) == Engine.AFFT
    assert recommended_gf16_engine(10, 4) == Engine.AFFT
    assert recommended_gf16_engine(6, 2) == Engine.AFFT
    assert recommended_gf16_engine(255, 1) == Engine.TOWER
    assert recommended_gf16_engine(255, 2) == Engine.AFFT
    assert recommended_gf16_engine(12, 4) == Engine.AFFT
    assert recommended_gf16_engine(13, 4) == Engine.TOWER


def test_engine_fields():
    assert Engine.STANDARD_CAUCHY.field() == Field.GF256
    assert Engine.GOOD_CAUCHY.field() == Field.GF256
    assert Engine.TOWER.field() == Field.GF65536
    assert Engine.AFFT.field() == Field.GF65536


def test_status_names():
    lib = _lib.load()
    assert lib.scrs_status_name(0) == b"Ok"
    assert lib.scrs_status_name(32) == b"DecodeError::DuplicateIndex"


def test_struct_layouts_match_the_header(tmp_path):
    # sizeof/offsetof from a C compiler reading include/*.h, against the
    # ctypes mirror the Python host uses (and a Rust #[repr(C)] would copy).
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {
        "scrs_status": _lib.Status, "scrs_profile": _lib.ProfileStruct,
        "scrs_push_outcome": _lib.PushOutcomeStruct, "scrs_job": _lib.JobStruct,
        "scrs_batch_config": _lib.BatchConfigStruct, "scrs_batch_stats": _lib.BatchStatsStruct,
        "scrs_server_config": _lib.ServerConfigStruct, "scrs_server_stats": _lib.ServerStatsStruct,
        "scrs_completion": _lib.CompletionStruct,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "scrs.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                 text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)
