"""Pins of the oracle's CASToR list-mode restatements (CPU), and the host side
of the list-mode drop-ins.

* ParseCASToRIDs: with no shift the LOR runs from the crystal with the smaller
  transaxial index; the TOF keeps its sign when that crystal is the event's
  first, and flips otherwise (equal indices: the ring of the larger ID comes
  first); with a shift the ring order follows the shifted indices while the
  detectors stay unshifted.
* ExportPrompts: events built at known cells and at TOF bin centres come back
  as exactly those counts; out-of-range TOFs are not counted but still set the
  randoms (word 0 * duration / bins) and normalisation (word 1) of their cell,
  where the file's last event wins; counts wrap at 2^16.
* InjectScatters: the kept events are those with |TOF| <= R and a cell, in
  file order, each with scale / (2R / B) * 2 / c / duration * scatter inserted
  after the attenuation word.
"""
import struct

import numpy as np
import pytest

from opensss_amd import castor
from opensss_amd.sinogram import SinogramCoordinates

N, NR = 32, 4


def _tables():
    lor, sino, _ = SinogramCoordinates(N, 1, 1, 1, 1, NR)
    return lor, sino


def _pack(flags, rows):
    fmt, _ = castor.GenerateListEventInfo(*flags)
    return b''.join(struct.pack(fmt, *r) for r in rows)


def test_oracle_event_format_matches_the_host_layout(oracle):
    for flags in [(0, 0, 0, 0, 1, 1), (1, 0, 1, 1, 1, 1), (0, 1, 1, 0, 1, 2), (1, 1, 1, 1, 1, 3)]:
        assert oracle.listmode_event_format(*flags) == castor.GenerateListEventInfo(*flags)
        fmt, _ = castor.GenerateListEventInfo(*flags)
        h = castor.listmode_header('a_df.Cdf', 1, attenuation=flags[0], scatter=flags[1], random=flags[2],
                                   normalization=flags[3], max_lines=flags[5])
        assert struct.calcsize(fmt) == 4 * castor.event_words(h)


@pytest.mark.parametrize('seed', [0, 1])
def test_parse_ids_closed_form_without_shift(oracle, seed):
    rng = np.random.default_rng(seed)
    _, pos = castor.GenerateListEventInfo(0, 0, 0, 0, 1, 1)
    for _ in range(500):
        ra, rb = rng.integers(0, NR, size=2)
        da, db = rng.integers(0, N, size=2)
        tof = float(np.float32(rng.uniform(-900, 900)))
        ida, idb = int(ra * N + da), int(rb * N + db)
        ev = (0, tof, ida, idb)
        got = oracle.parse_castor_ids(ev, pos, N, 0)
        if da < db:
            want = (da, ra, db, rb, tof)
        elif da > db:
            want = (db, rb, da, ra, -tof)
        else:
            # equal transaxial indices: the larger ID's ring first, TOF sign
            # flipped when the first crystal has the smaller ID
            want = (da, max(ra, rb), db, min(ra, rb), -tof if ida < idb else tof)
        assert got == want
        assert castor.ParseCASToRIDs(ev, pos, N, 0) == got


def test_parse_ids_shift_orders_rings_by_shifted_index(oracle):
    rng = np.random.default_rng(5)
    _, pos = castor.GenerateListEventInfo(0, 0, 0, 0, 1, 1)
    shift = 7
    for _ in range(500):
        ra, rb = rng.integers(0, NR, size=2)
        da, db = rng.integers(0, N, size=2)
        if da == db:
            continue
        ev = (0, 1.5, int(ra * N + da), int(rb * N + db))
        d1, r1, d2, r2, tof = oracle.parse_castor_ids(ev, pos, N, shift)
        assert (d1, d2) == (min(da, db), max(da, db))           # unshifted detectors
        sa, sb = (da - shift) % N, (db - shift) % N
        first_ring = ra if sa < sb else rb                       # ring of the smaller shifted index
        assert r1 == first_ring and r2 == (rb if first_ring == ra and sa < sb else ra)
        assert tof == (1.5 if sa < sb else -1.5)
        assert castor.ParseCASToRIDs(ev, pos, N, shift) == (d1, r1, d2, r2, tof)


def test_export_prompts_counts_events_built_at_known_cells(oracle):
    lor, sino = _tables()
    flags = (0, 0, 1, 1, 1, 1)     # randoms and normalisation words, TOF
    R, B, dur = 1000.0, 6, 300.0
    w = 2 * R / B
    rng = np.random.default_rng(2)
    rows, want = [], np.zeros((N + 1, N // 2, NR * NR, B), np.int64)
    last = {}
    for k in range(3000):
        ra, rb = rng.integers(0, NR, size=2)
        da, db = sorted(rng.choice(N, size=2, replace=False))
        b = int(rng.integers(0, B))
        out_of_range = k % 10 == 0
        tof = (R * 1.5) if out_of_range else (-R + (b + 0.5) * w)
        t = int(rng.integers(0, 1000))
        nv = float(np.float32(rng.uniform(0.5, 2)))
        rows.append((t, 0.25, nv, tof, int(ra * N + da), int(rb * N + db)))
        cell = (lor[da, db, 1], lor[da, db, 0], sino[ra, rb])
        if not out_of_range:
            want[cell + (b,)] += 1
        last[cell] = (t, nv)
    prompts, randoms, norm = oracle.export_prompts_loops(_pack(flags, rows), flags, dur, lor, sino, B, R)
    assert prompts.dtype == np.uint16 and np.array_equal(prompts, want)
    want_r = np.zeros(randoms.shape, np.float32)
    want_n = np.ones(norm.shape, np.float32)
    for cell, (t, nv) in last.items():
        want_r[cell] = np.float32(t * dur / B)
        want_n[cell] = np.float32(0.25)   # word 1 is the first float: the randoms correction here
    assert np.array_equal(randoms, want_r) and np.array_equal(norm, want_n)


def test_export_prompts_uint16_counts_wrap(oracle):
    lor, sino = _tables()
    flags = (0, 0, 0, 0, 1, 1)
    rows = [(0, 10.0, 3, 9)] * (65536 + 3)
    prompts, _, _ = oracle.export_prompts_loops(_pack(flags, rows), flags, 1.0, lor, sino, 4, 1000.0)
    assert prompts.sum() == 3


def test_inject_keeps_in_range_events_with_their_rate(oracle):
    lor, sino = _tables()
    flags = (1, 0, 0, 0, 1, 1)     # attenuation word, then the scatter rate goes after it
    R, B, dur = 1000.0, 4, 120.0
    w = 2 * R / B
    rng = np.random.default_rng(9)
    S = rng.uniform(0.1, 5.0, size=(B, N + 1, N // 2, NR * NR)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, size=B).astype(np.float32)
    rows, want = [], []
    factor = castor.injection_factors(scale, R, dur)
    for k in range(400):
        ra, rb = rng.integers(0, NR, size=2)
        da, db = sorted(rng.choice(N, size=2, replace=False))
        b = int(rng.integers(0, B))
        tof = 1.25 * R if k % 7 == 0 else -R + (b + 0.5) * w
        row = (k, 0.5, tof, int(ra * N + da), int(rb * N + db))
        rows.append(row)
        if k % 7:
            rate = np.float32(factor[b] * S[b, lor[da, db, 1], lor[da, db, 0], sino[ra, rb]])
            want.append(struct.pack('<1L3f2L', k, 0.5, rate, tof, row[3], row[4]))
    got = oracle.inject_scatters_loops(_pack(flags, rows), flags, dur, lor, sino, S, scale, R)
    assert got == b''.join(want)
    # the factor: scale / float32(2R/B) in float32, then * 2/c and / duration in float32
    f0 = np.float32(np.float32(scale[0] / np.float32(2 * R / B)) * np.float32(2 / 0.299792458))
    assert factor[0] == np.float32(f0 / np.float32(dur))


def test_header_round_trip_and_flag_quirk(tmp_path):
    h = castor.listmode_header('scan_df.Cdf', 12, random=True, max_lines=2, tof_range=750.0)
    castor.WriteHeaderFile(h, tmp_path / 'scan_df.Cdh')
    assert castor.ReadHeaderFile(tmp_path / 'scan_df.Cdh') == h
    text = (tmp_path / 'scan_df.Cdh').read_text() + 'Attenuation correction flag: 0\n'
    (tmp_path / 'q.Cdh').write_text(text)
    assert castor.ReadHeaderFile(tmp_path / 'q.Cdh')['attenuation_flag'] is True   # bool('0')


def test_datafile_size_check(tmp_path):
    h = castor.listmode_header('s_df.Cdf', 5)
    ev = castor.synthetic_events(h, 4, N, NR, np.random.default_rng(0))
    castor.write_listmode(tmp_path, 's_df.Cdh', h, ev)
    with pytest.raises(ValueError, match='not consistent'):
        castor.read_events(str(tmp_path), castor.ReadHeaderFile(tmp_path / 's_df.Cdh'))
