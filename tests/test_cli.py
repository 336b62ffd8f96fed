"""`ldsc l2` drop-in: formats and argument handling (CPU), and the whole
command on PLINK files against the CPU restatement (GPU)."""
import gzip
import os

import numpy as np
import pytest

from ldsc_amd import cli, synth


def test_rust_display_format():
    """Rust f64 Display used by format_m_vals (src/l2/io.rs:161-168)."""
    cases = {5.0: "5", 0.5: "0.5", 1e20: "100000000000000000000", 1e-7: "0.0000001",
             123.25: "123.25", -0.0: "-0", 0.1 + 0.2: "0.30000000000000004", 2500.0: "2500"}
    for v, s in cases.items():
        assert cli.rust_display(v) == s
    assert cli.format_m_vals([3.0, 0.75]) == "3\t0.75\n"


def test_inject_implicit_subcommand():
    assert cli.inject_implicit_subcommand(["--bfile", "x", "--l2", "--out", "o"]) == \
        ["--bfile", "x", "--out", "o"]
    assert cli.inject_implicit_subcommand(["l2", "--bfile", "x"]) == ["--bfile", "x"]


def test_read_annot_formats(tmp_path):
    full = tmp_path / "a.annot"
    full.write_text("CHR\tSNP\tBP\tCM\tbase\tcoding\n1\trs1\t10\t0\t1\t0\n1\trs2\t20\t0\t1\tNA\n")
    m, names = cli.read_annot_path(str(full), thin=False)
    assert names == ["base", "coding"] and m.tolist() == [[1, 0], [1, 0]]
    thin = tmp_path / "b.annot.gz"
    with gzip.open(thin, "wt") as f:
        f.write("x y\n0.5 2\n. 3\n")
    m, names = cli.read_annot_path(str(thin), thin=True)
    assert names == ["x", "y"] and m.tolist() == [[0.5, 2.0], [0.0, 3.0]]
    assert cli.resolve_annot_path(str(tmp_path / "b")) == str(thin)
    with pytest.raises(FileNotFoundError):
        cli.resolve_annot_path(str(tmp_path / "zz"))


def test_flag_conflicts():
    base = ["--bfile", "nope", "--out", "o"]
    for extra, msg in ((["--ld-wind-cm", "1", "--ld-wind-kb", "2"], "exactly one"),
                       (["--per-allele", "--pq-exp", "1"], "both --per-allele"),
                       (["--python-compat", "--snp-level-masking"], "conflicts"),
                       (["--sketch-maf-aware"], "requires --sketch"),
                       (["--chunk-size", "0"], "chunk_size must be >= 1")):
        with pytest.raises(SystemExit) as e:
            cli.run(cli.build_parser().parse_args(base + extra))
        assert msg in str(e.value)


def _read_ldscore(path):
    with gzip.open(path, "rt") as f:
        rows = [ln.rstrip("\n").split("\t") for ln in f]
    return rows[0], rows[1:]


@pytest.mark.gpu
def test_cli_end_to_end(tmp_path, oracle):
    p = synth.make_panel(n=350, m=900, seed=17, miss_rate=0.02)
    p.chr[:] = np.where(np.arange(p.m) < 500, 1, 2)
    prefix = str(tmp_path / "panel")
    synth.write_plink(prefix, p)
    annot = synth.make_annot(p.m, seed=3, n_bin=2, n_cont=1)
    for c in (1, 2):
        sel = np.where(p.chr == c)[0]
        with open(f"{prefix}{c}.annot", "w") as f:
            f.write("CHR\tSNP\tBP\tCM\tA\tB\tC\n")
            for j in sel:
                f.write(f"{c}\trs{j + 1}\t{int(p.bp[j])}\t{float(p.cm[j])!r}\t" +
                        "\t".join(repr(float(v)) for v in annot[j]) + "\n")
    out = str(tmp_path / "out")
    rc = cli.main(["--bfile", prefix, "--l2", "--out", out, "--ld-wind-cm", "0.2",
                   "--annot", prefix, "--snp-level-masking", "--maf", "0.02", "--yes-really"])
    assert rc == 0
    # oracle on the same SNPs: MAF prefilter (het + missing < N, maf > 0.02)
    st = oracle.stats(p.bed_bytes(), p.n, np.arange(p.m))
    cnt = st[:, 1].astype(float)
    freq = st[:, 0] / (2 * cnt)
    maf_all = np.minimum(freq, 1 - freq)
    keep = ((st[:, 3] + (p.n - st[:, 1])) < p.n) & (maf_all > 0.02)
    sel = np.where(keep)[0]
    exp, emaf = oracle.l2(p.bed_bytes(), p.n, p.chr[sel], p.cm[sel], p.bp[sel], bed_idx=sel,
                          mode=0, window=0.2, chunk=200, masking=True, annot=annot[sel],
                          per_chromosome=True)
    header, rows = _read_ldscore(out + ".l2.ldscore.gz")
    assert header == ["CHR", "SNP", "BP", "AL2", "BL2", "CL2"]
    printed = np.where(emaf > 0.02)[0]
    assert [r[1] for r in rows] == [f"rs{sel[i] + 1}" for i in printed]
    got = np.array([[float(x) for x in r[3:]] for r in rows])
    assert np.max(np.abs(got - exp[printed])) <= 0.0005 + 1e-9
    m_line = open(out + ".l2.M").read()
    assert m_line == cli.format_m_vals(np.cumsum(annot[sel], axis=0)[-1])
    m5 = open(out + ".l2.M_5_50").read()
    assert m5 == cli.format_m_vals(np.cumsum(annot[sel][emaf > 0.05], axis=0)[-1])
    for c in (1, 2):
        assert os.path.exists(f"{out}{c}.l2.ldscore.gz") and os.path.exists(f"{out}{c}.l2.M")


class _OraclePanel:
    """Stands in for the device panel so the CLI's host logic can be checked
    on the CPU with the restatement doing the arithmetic (tests only)."""

    def __init__(self, bed, n, m, keep=None):
        self.bed, self.n_fam, self.m, self.keep = bed, n, m, keep
        self.n_indiv = n if keep is None else len(keep)

    def close(self):
        pass


def _oracle_backend(monkeypatch, oracle):
    def decode(panel, idx, **kw):
        return None, None, oracle.stats(panel.bed, panel.n_fam, idx, keep=panel.keep), None

    def compute(panel, snps, cfg, annot=None):
        l2, maf = oracle.l2(panel.bed, panel.n_fam, snps["chr"], snps["cm"], snps["bp"],
                            bed_idx=snps["bed_idx"], mode=cfg.mode.kind, window=cfg.mode.value,
                            chunk=cfg.chunk_size, masking=cfg.snp_level_masking, annot=annot,
                            per_chromosome=cfg.per_chromosome, keep=panel.keep,
                            pq_exp=cfg.pq_exp, use_f32=cfg.use_f32)
        return l2, maf, None

    monkeypatch.setattr(cli, "Panel", _OraclePanel)
    monkeypatch.setattr(cli, "decode", decode)
    monkeypatch.setattr(cli, "compute_ldscore", compute)


def test_cli_host_logic_with_oracle(tmp_path, monkeypatch, oracle):
    """--extract, --keep, --print-snps, --per-allele (pq pre-scaling of a thin
    annotation after the MAF prefilter), --python-compat chunking."""
    _oracle_backend(monkeypatch, oracle)
    p = synth.make_panel(n=120, m=300, seed=23, miss_rate=0.03)
    p.chr[:] = np.where(np.arange(p.m) < 170, 3, 5)
    prefix = str(tmp_path / "p")
    synth.write_plink(prefix, p)
    ext = np.sort(np.random.default_rng(0).choice(p.m, 250, replace=False))
    (tmp_path / "ext.txt").write_text("".join(f"rs{j + 1}\n" for j in ext))
    keep = np.arange(0, p.n, 2)
    (tmp_path / "keep.txt").write_text("".join(f"F{i}\tI{i}\n" for i in keep))
    pr = ext[::3]
    (tmp_path / "print.txt").write_text("".join(f"rs{j + 1}\n" for j in pr))
    a = np.random.default_rng(1).random((250, 2))
    with open(tmp_path / "ann.annot", "w") as f:
        f.write("X\tY\n" + "".join(f"{float(u)!r}\t{float(v)!r}\n" for u, v in a))
    out = str(tmp_path / "o")
    rc = cli.main(["l2", "--bfile", prefix, "--out", out, "--ld-wind-kb", "40",
                   "--extract", str(tmp_path / "ext.txt"), "--keep", str(tmp_path / "keep.txt"),
                   "--print-snps", str(tmp_path / "print.txt"), "--per-allele",
                   "--annot", str(tmp_path / "ann.annot"), "--thin-annot", "--python-compat"])
    assert rc == 0
    st = oracle.stats(p.bed_bytes(), p.n, ext, keep=keep)
    cnt = st[:, 1].astype(float)
    freq = st[:, 0] / (2 * cnt)
    maf = np.minimum(freq, 1 - freq)
    ok = ((st[:, 3] + (len(keep) - st[:, 1])) < len(keep)) & (maf > 0.0)
    sel = ext[ok]
    aa = a[ok] * (maf[ok] * (1 - maf[ok]))[:, None]
    exp, emaf = oracle.l2(p.bed_bytes(), p.n, p.chr[sel], p.cm[sel], p.bp[sel], bed_idx=sel,
                          mode=1, window=40.0, chunk=50, annot=aa, keep=keep)
    header, rows = _read_ldscore(out + ".l2.ldscore.gz")
    assert header == ["CHR", "SNP", "BP", "XL2_S1", "YL2_S1"]
    shown = [i for i, j in enumerate(sel) if j in set(pr)]
    assert [r[1] for r in rows] == [f"rs{sel[i] + 1}" for i in shown]
    got = np.array([[float(x) for x in r[3:]] for r in rows])
    assert np.max(np.abs(got - exp[shown])) <= 0.0005 + 1e-12
    assert open(out + ".l2.M").read() == cli.format_m_vals(np.cumsum(aa, axis=0)[-1])
    assert sorted(os.path.basename(x) for x in os.listdir(tmp_path) if x.startswith("o")) == \
        sorted(["o.l2.ldscore.gz", "o.l2.M", "o.l2.M_5_50", "o3.l2.ldscore.gz", "o3.l2.M",
                "o3.l2.M_5_50", "o5.l2.ldscore.gz", "o5.l2.M", "o5.l2.M_5_50"])
