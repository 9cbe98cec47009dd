"""The C ABI library loads and exports every symbol include/culda.h declares;
the ctypes mirror lays out every ABI struct as a C compiler does; the CLI
parses the reference flags.  No compute calls (no GPU needed)."""
import re
import subprocess
from pathlib import Path

import pytest

import culda_cgs_amd as cg

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "culda.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(culda_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = cg.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared function is bound with a signature in the Python layer
    from culda_cgs_amd import _SIGS
    assert set(names) <= set(_SIGS), set(names) - set(_SIGS)
    assert L.culda_abi_version() == 1


def test_error_paths_without_a_device():
    with pytest.raises(cg.CuldaError) as e:  # k not a multiple of 16
        cg.Sampler(k=1000)
    assert e.value.status == -1
    with pytest.raises(cg.CuldaError):
        cg.Sampler(k=64, alpha=-1.0)
    assert "alpha" in cg.lib().culda_last_error().decode()


def run_cli(*args):
    return subprocess.run([str(cg.CLI_PATH), *args], capture_output=True, text=True, timeout=60)


def test_cli_flag_errors():
    r = run_cli()
    assert r.returncode == 1 and "No argument found" in r.stdout
    r = run_cli("-k", "abc", "-i", "x")
    assert r.returncode == 1 and "-k should be followed by a positive integer" in r.stdout
    r = run_cli("-t")
    assert r.returncode == 1 and "need to specify" in r.stdout
    r = run_cli("-k", "1024", "--bogus")
    assert r.returncode == 1 and "unknown option" in r.stdout
    r = run_cli("-a", "50/0", "-i", "x")
    assert r.returncode == 1


def test_cli_missing_input_reports_and_exits(tmp_path):
    r = run_cli("-k", "1000", "-a", "50/1024", "-i", str(tmp_path / "none"))
    assert "rounded to multiples of 32" in r.stdout
    assert "alpha       :0.048828" in r.stdout  # 50/1024, not atof's 50
    assert r.returncode == 2 and "open failed" in r.stderr


def test_check_corpus_limits_without_a_device():
    """culda_check_corpus (run by culda_upload before anything is allocated)
    refuses k outside the sampler's range and a doc whose theta row would not
    fit the sampler's LDS rows (30,576 entries: a doc of ~1.93M tokens at
    k = 16); an ordinary corpus passes."""
    import numpy as np
    from helpers import chunks_from_docs, product_corpus
    c = cg.Corpus.generate(300, 50, 5, 1, len_mean=30.0, seed=3, num_threads=2)
    for k in (16, 64, 1024):
        c.check(k)
    for k in (0, 8, 1000, 2048):
        with pytest.raises(cg.CuldaError) as e:
            c.check(k)
        assert e.value.status == -1
    rng = np.random.default_rng(0)
    V = 50
    big = product_corpus(V, *chunks_from_docs(V, [rng.integers(0, V, 1_950_000), rng.integers(0, V, 10)], 1))
    with pytest.raises(cg.CuldaError) as e:
        big.check(16)
    assert e.value.status == -7 and "LDS" in str(e.value) and "doc 0" in str(e.value)
    ok = product_corpus(V, *chunks_from_docs(V, [rng.integers(0, V, 1_800_000), rng.integers(0, V, 10)], 1))
    ok.check(16)


# every struct of the ABI and its ctypes mirror
ABI_STRUCTS = [("culda_corpus_info", cg.CorpusInfo), ("culda_chunk_view", cg.ChunkView),
               ("culda_gen_params", cg.GenParams), ("culda_params", cg.Params), ("culda_stats", cg.Stats),
               ("culda_mh_params", cg.MHParams), ("culda_tune", cg.Tune)]


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and every field offset, as a C compiler lays out include/culda.h,
    equal the ctypes mirror's (a field name the header lacks fails to compile)."""
    import ctypes
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "culda.h"', "int main(void) {"]
    for cname, cls in ABI_STRUCTS:
        lines.append(f'printf("{cname} . %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {(a, b): int(c) for a, b, c in (line.split() for line in out.splitlines())}
    for cname, cls in ABI_STRUCTS:
        assert got[(cname, ".")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, f"{cname}.{f}"
