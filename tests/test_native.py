"""Native host checks of code shared with the device kernels (CPU only)."""
import os
import subprocess

from conftest import PKG, REPO


def test_moddiv_matches_operator(tmp_path):
    exe = str(tmp_path / "moddiv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(PKG, "csrc"),
                    os.path.join(REPO, "tests", "native", "moddiv_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "moddiv ok" in out.stdout


def test_exact_pairs_and_verify_reject(tmp_path):
    exe = str(tmp_path / "pairs_check")
    csrc = os.path.join(PKG, "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", csrc,
                    os.path.join(REPO, "tests", "native", "pairs_check.cpp"), os.path.join(csrc, "fc5_host.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "pairs ok" in out.stdout


def test_solution_record_fields(tmp_path):
    """the SOLUTION record of a verified identity: sorted terms, refusals, gcd"""
    exe = str(tmp_path / "solution_check")
    csrc = os.path.join(PKG, "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", csrc,
                    os.path.join(REPO, "tests", "native", "solution_check.cpp"), os.path.join(csrc, "fc5_host.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "solution ok" in out.stdout


def test_sieve_words_equal_reference_gate(tmp_path):
    """the class-5 bit sieve (five small moduli, 32 grid points a word) is the reference gate"""
    exe = str(tmp_path / "sieve_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(PKG, "csrc"),
                    os.path.join(REPO, "tests", "native", "sieve_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "sieve ok" in out.stdout
