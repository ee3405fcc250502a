"""Autotuner table (no GPU needed): variant counts, the kernel a shape
selects, and the JSON cache round trip."""
import os

import pytest


@pytest.fixture()
def lc():
    import libconv
    libconv.tune_clear()
    yield libconv
    libconv.tune_clear()


K256 = dict(n=[256] * 3, bc=-1, nx=[284] * 3, ny=[256] * 3)


def test_variant_counts(lc):
    assert lc.variant_count(lc.FORM_SEP) == 73
    assert lc.variant_count(lc.FORM_SUM) == 57
    assert lc.variant_count(lc.FORM_DWT) == 18
    assert lc.variant_count(lc.FORM_IDWT) == 18


def test_default_kernel_names(lc):
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **K256) == 'sum3d2w_fused<double, lc::FD2, 32, 16, 8, 8, 2, 4, 0, 2>'
    assert lc.kernel_name(4, lc.FORM_SUM, lc.FIL_D2, **K256) == 'sum3d2_fused<float, lc::FD2, 32, 16, 8, 8, 2, 2, 0>'
    assert lc.kernel_name(4, lc.FORM_SEP, lc.FIL_MD, [284] * 3, -1, [299] * 3, [284] * 3) == \
        'sepq_fused<float, lc::FMD, 32, 32, 8, 4, 2, 2>'
    assert lc.kernel_name(8, lc.FORM_SEP, lc.FIL_MD, [284] * 3, -1, [299] * 3, [284] * 3) == \
        'sepq_fused<double, lc::FMD, 32, 16, 4, 2, 2, 2>'
    assert lc.kernel_name(8, lc.FORM_IDWT, lc.WAV_SYM8, [96] * 3, 0, [96] * 3, [96] * 3).startswith(
        'wavq_fused<double, lc::Idwt<')
    assert lc.kernel_name(8, lc.FORM_IDWT, lc.WAV_SYM8, [96] * 3, 0, [96] * 3, [96] * 3).endswith(
        ', 8, 8, 2, 2, 1, 0>')
    assert lc.kernel_name(4, lc.FORM_DWT, lc.WAV_SYM8, [96] * 3, 0, [96] * 3, [96] * 3).endswith(
        ', 16, 8, 4, 2, 1, 0>')


def test_planes_beyond_buffer_records_use_pointer_kernels(lc):
    # the two-plane and pipelined kernels read planes through 32-bit buffer
    # records; a plane of more than 2^31 - 1 bytes selects the single-plane
    # kernels instead
    big = dict(n=[16400, 16400, 8], bc=0, nx=[16400, 16400, 8], ny=[16400, 16400, 8])
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **big).startswith('sum3d_fused<double')
    assert lc.kernel_name(8, lc.FORM_SEP, lc.FIL_MD, **big).startswith('sep3d_fused<double')
    assert lc.kernel_name(4, lc.FORM_SEP, lc.FIL_MD, **big).startswith('sepq_fused<float')


def test_environment_selects_variant(lc, monkeypatch):
    monkeypatch.setenv('LIBCONV_SUM3D_VARIANT', '6')
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **K256) == 'sum3d_ws<double, lc::FD2, 64, 8, 4, 4, 2>'
    monkeypatch.setenv('LIBCONV_SUM3D_VARIANT', '99')   # out of range: default
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **K256).endswith('32, 16, 8, 8, 2, 4, 0, 2>')


def test_json_cache_round_trip(lc, tmp_path, monkeypatch):
    monkeypatch.setenv('LIBCONV_SUM3D_VARIANT', '6')     # the table wins over the environment
    p = tmp_path / 'tune.json'
    p.write_text('{"libconv_tune": 1, "entries": [\n'
                 ' {"prec": 8, "form": 1, "op": 2, "n": [256, 256, 256], "bc": [-1, -1, -1], '
                 '"nx": [284, 284, 284], "ny": [256, 256, 256], "variant": 3, "ms": 0.25},\n'
                 ' {"prec": 8, "form": 2, "op": 0, "n": [96, 96, 96], "bc": [0, 0, 0], '
                 '"nx": [96, 96, 96], "ny": [96, 96, 96], "variant": 2, "ms": 0.1},\n'
                 ' {"prec": 8, "form": 1, "op": 2, "n": [8, 8, 8], "bc": [0, 0, 0], '
                 '"nx": [8, 8, 8], "ny": [8, 8, 8], "variant": 99, "ms": 0.1}\n]}\n')
    lc.tune_load(str(p))
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **K256) == 'sum3d_fused<double, lc::FD2, 64, 8, 2, 4, 4, 2, 1>'
    assert ', 32, 4, 4, 2, 2, 1>' in lc.kernel_name(8, lc.FORM_DWT, lc.WAV_SYM8, [96] * 3, 0, [96] * 3, [96] * 3)
    # the entry with an unknown variant is ignored; other precisions are separate
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, [8] * 3, 0, [8] * 3, [8] * 3).startswith('sum3d_ws<')
    assert lc.kernel_name(4, lc.FORM_SUM, lc.FIL_D2, **K256).startswith('sum3d_ws<float')
    q = tmp_path / 'again.json'
    lc.tune_save(str(q))
    text = q.read_text()
    assert text.count('"variant"') == 2 and '"kernel": "sum3d_fused<double, lc::FD2, 64, 8, 2, 4, 4, 2, 1>"' in text
    lc.tune_clear()
    lc.tune_load(str(q))
    assert lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, **K256).endswith('64, 8, 2, 4, 4, 2, 1>')


def test_tuner_rejects_bad_arguments(lc, tmp_path):
    with pytest.raises(lc.LibconvError):
        lc.kernel_name(8, 7, 0, [8] * 3, 0, [8] * 3, [8] * 3)
    with pytest.raises(lc.LibconvError):
        lc.kernel_name(8, lc.FORM_DWT, lc.WAV_SYMICOMB8, [8] * 3, 0, [8] * 3, [8] * 3)
    with pytest.raises(lc.LibconvError):
        lc.kernel_name(8, lc.FORM_SUM, lc.FIL_D2, [8] * 3, -1, [8] * 3, [8] * 3)  # shrink input short
    with pytest.raises(lc.LibconvError):
        lc.tune_load(str(tmp_path / 'missing.json'))
