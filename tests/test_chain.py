"""chain_add (softwarerenderer_amd/csrc/swr_chain.h), the closed form the Span raster uses for
the reference's per-pixel addition chains (PixelData::stepX along drawSpan,
PixelShaderBase.h:96-112), against the addition loop itself, compiled for the host: two
million cases over magnitudes, ties, zero crossings, binade edges, subnormals and
non-finite values must agree bit for bit."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_chain_add_matches_the_addition_loop():
    tmp = tempfile.mkdtemp(prefix="swr_chain_")
    exe = os.path.join(tmp, "chain_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__host__=",
                    "-D__device__=", "-I" + os.path.join(ROOT, "softwarerenderer_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "chain_check.cpp"), "-o", exe],
                   check=True)
    out = subprocess.run([exe, "2000000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "mismatches 0" in out.stdout
