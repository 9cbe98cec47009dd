"""The C ABI library loads and exports every symbol include/culda.h declares;
the CLI parses the reference flags.  No compute calls (no GPU needed)."""
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
