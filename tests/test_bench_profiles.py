"""bench.py's counter figures from profiles/: read only while the kernel sources
match the fingerprint the summary records (host logic, no GPU)."""

import bench


def _summary(path, fp, value=4.2e9):
    path.write_text(f"# header\ntraffic_bytes_per_launch[fft_row] = {value:.4e}  (kernel)\n"
                    f"sources_sha256 = {fp}  (files)\n")
    return str(path)


def test_figure_read_when_sources_match(tmp_path):
    fp = bench.source_fingerprint(bench.FFT_SOURCES)
    f = _summary(tmp_path / "s.txt", fp)
    v, note = bench.profile_figure(f, bench.FFT_SOURCES, "traffic_bytes_per_launch[fft_row]")
    assert v == 4.2e9 and note is None


def test_figure_dropped_when_sources_differ(tmp_path):
    f = _summary(tmp_path / "s.txt", "0123456789abcdef")
    v, note = bench.profile_figure(f, bench.FFT_SOURCES, "traffic_bytes_per_launch[fft_row]")
    assert v is None and "other kernel sources" in note


def test_missing_summary_and_key(tmp_path):
    v, note = bench.profile_figure(str(tmp_path / "none.txt"), bench.FFT_SOURCES, "x")
    assert v is None and "missing" in note
    fp = bench.source_fingerprint(bench.FFT_SOURCES)
    f = _summary(tmp_path / "s.txt", fp)
    v, note = bench.profile_figure(f, bench.FFT_SOURCES, "traffic_bytes_per_launch[fft_half]")
    assert v is None and "not in" in note


def test_fingerprint_covers_each_source():
    """Each listed file changes the fingerprint (a dropped file would let a stale
    summary through)."""
    full = bench.source_fingerprint(bench.FFT_SOURCES)
    for i in range(len(bench.FFT_SOURCES)):
        assert bench.source_fingerprint(bench.FFT_SOURCES[:i] + bench.FFT_SOURCES[i + 1:]) != full
