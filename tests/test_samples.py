"""Sample alignment before the path (src/main.rs:291-571, src/bed.rs:456-478):
the reference's own covariate known-answer test (src/main.rs:626-690), its
error texts, intersection, and the estimation-panel alignment."""
import numpy as np
import pytest

from schnellfmm_amd import samples as SM


def _fam(iids, fid="FAM"):
    return [SM.FamRecord(fid, i, "0", "0", 0, "-9") for i in iids]


def test_covariate_imputed_values_correct(tmp_path):
    # src/main.rs:626-690: '.', 'NA' and empty cells take the observed column mean
    p = tmp_path / "cov.tsv"
    p.write_text("sample_id\tenv_0\tenv_1\n"
                 "IND0\t.\t0.5\n"
                 "IND1\t1.0\tNA\n"
                 "IND2\t4.0\t\n"
                 "IND3\t5.0\t3.5\n"
                 "IND4\t6.0\t4.5\n")
    names, x, kept = SM.load_covariates(str(p), _fam([f"IND{i}" for i in range(5)]))
    assert names == ["env_0", "env_1"]
    assert kept == [0, 1, 2, 3, 4]
    assert x.shape == (5, 2)
    m0 = (1.0 + 4.0 + 5.0 + 6.0) / 4.0
    m1 = (0.5 + 3.5 + 4.5) / 3.0
    assert abs(x[0, 0] - m0) < 1e-12
    assert np.allclose(x[1:, 0], [1.0, 4.0, 5.0, 6.0], atol=1e-12, rtol=0)
    assert abs(x[0, 1] - 0.5) < 1e-12 and abs(x[1, 1] - m1) < 1e-12 and abs(x[2, 1] - m1) < 1e-12
    assert abs(x[3, 1] - 3.5) < 1e-12 and abs(x[4, 1] - 4.5) < 1e-12


def test_covariates_fam_order_csv_and_intersect(tmp_path):
    p = tmp_path / "cov.csv"
    p.write_text(",a,b\nS3,3,30\nS1,1,10\nS9,9,90\nS2,2,20\n")
    fam = _fam(["S1", "S2", "S3", "S4"])
    with pytest.raises(ValueError, match=r"Sample 'S4' \(FID='FAM'\) from .fam file not found in covariate file"):
        SM.load_covariates(str(p), fam)
    names, x, kept = SM.load_covariates(str(p), fam, intersect=True)
    assert names == ["a", "b"] and kept == [0, 1, 2]
    assert np.array_equal(x, [[1, 10], [2, 20], [3, 30]])
    with pytest.raises(ValueError, match=r"No samples in common between .fam file \(1 samples\)"):
        SM.load_covariates(str(p), _fam(["Z"]), intersect=True)


@pytest.mark.parametrize("text,msg", [
    ("id\n", r"header has only 1 field\(s\)"),
    ("", "Covariate file is empty"),
    ("id\ta\n", "Covariate file has no data rows"),
    ("id\ta\tb\nS1\t1\n", r"Line 2 has 2 fields, expected 3 \(1 sample ID \+ 2 covariates\)"),
    ("id\ta\nS1\tx1\n", r"Failed to parse value 'x1' for covariate 'a' on line 2 \(sample 'S1'\)"),
    ("id\ta\nS1\t1_0\n", r"Failed to parse value '1_0'"),
    ("id\ta\nS1\tinf\n", r"Infinite value \(inf\) for covariate 'a' on line 2"),
    ("id\ta\nS1\t1\n\nS1\t2\n", r"Duplicate sample ID 'S1' in covariate file \(line 4\)"),
])
def test_covariate_errors(tmp_path, text, msg):
    p = tmp_path / "c.tsv"
    p.write_text(text)
    with pytest.raises(ValueError, match=msg):
        SM.load_covariates(str(p), _fam(["S1"]))


def test_read_fam_and_errors(tmp_path):
    p = tmp_path / "x.fam"
    p.write_text("F1 I1 0 0 1 -9\nF2\tI2\t0\t0\t2\t1.5\r\n")
    recs = SM.read_fam(str(p))
    assert [(r.fid, r.iid, r.sex, r.pheno) for r in recs] == [("F1", "I1", 1, "-9"), ("F2", "I2", 2, "1.5")]
    p.write_text("F1 I1 0 0 x -9\n")
    with pytest.raises(ValueError, match="FAM line 1: invalid sex 'x'"):
        SM.read_fam(str(p))
    p.write_text("F1 I1 0 0 1\n")
    with pytest.raises(ValueError, match="FAM line 1 has 5 fields, expected 6"):
        SM.read_fam(str(p))


def test_align_est_bed_samples():
    main = _fam(["A", "B", "C"])
    assert SM.align_est_bed_samples(_fam(["A", "B", "C"]), main) is None
    assert SM.align_est_bed_samples(_fam(["C", "A", "X", "B"]), main) == [1, 3, 0]
    with pytest.raises(ValueError, match=r"est-bed is missing 1 sample\(s\) that are in the main BED file: \[C\]"):
        SM.align_est_bed_samples(_fam(["A", "B"]), main)
    many = _fam([f"M{i}" for i in range(7)])
    # the reference closes the bracket before the ellipsis (src/main.rs:314-319)
    with pytest.raises(ValueError, match=r"missing 7 sample\(s\).*\[M0, M1, M2, M3, M4\], \.\.\.$"):
        SM.align_est_bed_samples(_fam(["A"]), many)
