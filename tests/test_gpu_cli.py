"""The command-line search (C++ host API over the C ABI) against the oracle:
bit-vector files in, reference result text out."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from oracle import records as R

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "sbwt_gpu_amd", "sbwt_search")
ACGT = np.array(list("ACGT"))


@pytest.mark.parametrize("extra", [[], ["--independent"], ["--without-presearch"], ["--n-as-a"], ["--layout=32"],
                                   ["--pipelined", "--batch-chars=2000"], ["--pipelined", "--independent"],
                                   ["--output=nodes"], ["--output=bitmap", "--batch-chars=2000"]])
def test_cli_matches_oracle_text(tmp_path, extra):
    rng = np.random.default_rng(8)
    g = "".join(ACGT[rng.integers(0, 4, 50_000)])
    k = 27
    s = O.SBWT([g], k)
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], s.rows[c], s.n)
    O.bv_write(str(tmp_path / "grp.bin"), s.group, s.n)
    lines = [">header"]
    for i in range(300):
        p = int(rng.integers(0, len(g) - 200))
        r = np.array(list(g[p:p + int(rng.integers(10, 200))]))
        r[rng.random(r.size) < 0.01] = "N"
        lines.append("".join(r) if i % 5 else "".join(ACGT[rng.integers(0, 4, 80)]))
    text = ("\n".join(lines) + "\n").encode()
    q = tmp_path / "q.txt"
    q.write_bytes(text)
    args = [CLI, f"--k={k}", "--k-presearch=9", *[f"--file-{c}={p}" for c, p in zip("acgt", paths)],
            f"--file-groups={tmp_path / 'grp.bin'}", f"--file-queries={q}", *extra]
    out = subprocess.run(args, capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()
    oix = O.Index(s.rows, s.n)
    oix.presearch(0 if "--without-presearch" in extra else 9)
    chars, rs = O.parse_lines(text)
    want = O.search_reads(oix, chars, rs, k, n_breaks="--n-as-a" not in extra)
    if "--output=nodes" in extra or "--output=bitmap" in extra:
        from sbwt_gpu_amd import read_binary_results

        dec = read_binary_results(out.stdout)
        np.testing.assert_array_equal(dec["row_kmers"], np.diff(O.kmer_offsets(rs, k)))
        np.testing.assert_array_equal(dec["values"], want if dec["kind"] == "nodes" else want != np.uint64(0xFFFFFFFFFFFFFFFF))
        return
    assert out.stdout == O.format_results(want, O.kmer_offsets(rs, k))


def test_cli_errors(tmp_path):
    r = subprocess.run([CLI, "--k=31", f"--file-a={tmp_path}/nope", f"--file-c={tmp_path}/nope",
                        f"--file-g={tmp_path}/nope", f"--file-t={tmp_path}/nope", "--file-queries=x"],
                       capture_output=True, timeout=60)
    assert r.returncode == 1 and b"can not be opened" in r.stderr


@pytest.mark.parametrize("fmt", ["fasta", "fastq"])
def test_cli_records_match_oracle(tmp_path, fmt):
    """--records: one row per FASTA / FASTQ record (wrapped lines, CRLF),
    searched and printed like the oracle's search of the records."""
    rng = np.random.default_rng(9)
    g = "".join(ACGT[rng.integers(0, 4, 40_000)])
    k = 25
    s = O.SBWT([g], k)
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], s.rows[c], s.n)
    parts = []
    for i in range(200):
        p = int(rng.integers(0, len(g) - 300))
        seq = g[p:p + int(rng.integers(5, 300))]
        wrap = [seq[j:j + 60] for j in range(0, len(seq), 60)]
        if fmt == "fasta":
            parts += [f">r{i}"] + wrap
        else:
            parts += [f"@r{i}"] + wrap + ["+"] + ["I" * len(w) for w in wrap]
    text = ("\r\n".join(parts) + "\r\n").encode()
    q = tmp_path / "q.txt"
    q.write_bytes(text)
    args = [CLI, f"--k={k}", "--k-presearch=8", *[f"--file-{c}={p}" for c, p in zip("acgt", paths)],
            f"--file-queries={q}", "--records", "--batch-chars=3000"]
    out = subprocess.run(args, capture_output=True, timeout=120)
    assert out.returncode == 0, out.stderr.decode()
    oix = O.Index(s.rows, s.n)
    oix.presearch(8)
    chars, rs = R.parse_records(text)
    want = O.search_reads(oix, chars, rs, k)
    assert out.stdout == O.format_results(want, O.kmer_offsets(rs, k))


def test_cli_index_image(tmp_path):
    """--save-image checkpoints the index built from the four files;
    --file-image resumes from it alone (k comes from the image) and prints
    the same text."""
    rng = np.random.default_rng(12)
    g = "".join(ACGT[rng.integers(0, 4, 30_000)])
    k = 23
    s = O.SBWT([g], k)
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], s.rows[c], s.n)
    O.bv_write(str(tmp_path / "grp.bin"), s.group, s.n)
    lines = [g[p:p + 120] for p in rng.integers(0, len(g) - 120, 200)] + ["".join(ACGT[rng.integers(0, 4, 90)])]
    q = tmp_path / "q.txt"
    q.write_bytes(("\n".join(lines) + "\n").encode())
    img = tmp_path / "ix.img"
    first = subprocess.run([CLI, f"--k={k}", "--k-presearch=8", *[f"--file-{c}={p}" for c, p in zip("acgt", paths)],
                            f"--file-groups={tmp_path / 'grp.bin'}", f"--file-queries={q}", f"--save-image={img}"],
                           capture_output=True, timeout=120)
    assert first.returncode == 0, first.stderr.decode()
    second = subprocess.run([CLI, f"--file-image={img}", "--k-presearch=8", f"--file-queries={q}"],
                            capture_output=True, timeout=120)
    assert second.returncode == 0, second.stderr.decode()
    assert second.stdout == first.stdout and len(first.stdout) > 0
    oix = O.Index(s.rows, s.n)
    oix.presearch(8)
    chars, rs = O.parse_lines(q.read_bytes())
    assert first.stdout == O.format_results(O.search_reads(oix, chars, rs, k), O.kmer_offsets(rs, k))
