"""GPU list-mode histogramming and scatter injection against the oracle's
restatements of the reference loops (Timber/Datafile.py), bit for bit: the
uint16 prompts, the float32 randoms and normalisation, and the bytes of the
injected datafile.  Events are random crystal pairs over the whole scanner
(TOF over +-1.2 times the range), on layouts with and without correction
words and with two lines per event, with mash 1 and 2, a crystal shift, and a
span ring table whose excluded pairs (-1) drop their events."""
import numpy as np
import pytest

from opensss_amd import castor
from opensss_amd.data_reduction import SinogramLUT
from opensss_amd.sinogram import SinogramCoordinates

pytestmark = pytest.mark.gpu

NR = 6


def _scanner(n_crystals, mash):
    lor, sino, _ = SinogramCoordinates(1, 1, 1, 1, n_crystals // mash, NR, ID_Shidt=False)
    return lor, sino


CASES = [
    # (flags: attenuation, scatter, random, norm, max_lines), mash, shift, span
    ((0, 0, 1, 1, 1), 1, 0, False),
    ((1, 0, 0, 0, 1), 2, 3, False),
    ((1, 0, 1, 1, 2), 1, 5, True),
    ((0, 0, 0, 1, 1), 2, 0, True),
]


@pytest.mark.parametrize('flags,mash,shift,span', CASES)
def test_histogram_and_injection_match_the_reference_loops(oracle, flags, mash, shift, span):
    n_crystals = 64
    lor, sino = _scanner(n_crystals, mash)
    table = SinogramLUT(NR, 3, 3)[2] if span else sino
    h = castor.listmode_header('t_df.Cdf', 0, duration=321.0, tof_range=800.0, attenuation=flags[0],
                               scatter=flags[1], random=flags[2], normalization=flags[3], max_lines=flags[4])
    rng = np.random.default_rng(hash((flags, mash, shift, span)) % 2 ** 32)
    ev = castor.synthetic_events(h, 40000, n_crystals, NR, rng)
    idw = castor.GenerateEventClass(h)[1][4]
    ev[:3000, idw:idw + 2] = ev[0, idw:idw + 2]   # one crystal pair in many events: a hot cell
    B = 5
    ofl = (flags[0], flags[1], flags[2], flags[3], 1, flags[4])
    data = ev.astype('<u4').tobytes()
    prompts, randoms, norm = castor.histogram_events(ev, h, lor, table, B, h['tof_range'], mash, shift)
    want = oracle.export_prompts_loops(data, ofl, h['duration'], lor, table, B, h['tof_range'], mash, shift)
    assert np.array_equal(prompts, want[0])
    assert prompts.sum() > 0
    for g, w_ in ((randoms, want[1]), (norm, want[2])):
        assert (g is None) == (w_ is None)
        if g is not None:
            assert np.array_equal(g, w_)
    # injection with random scatter sinograms and scale factors
    S = rng.uniform(0.1, 3.0, size=(B,) + prompts.shape[:3]).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, size=B).astype(np.float32)
    f = castor.injection_factors(scale, h['tof_range'], h['duration'])
    kept = castor.inject_events(ev, h, lor, table, S, f, h['tof_range'], mash, shift)
    want_bytes = oracle.inject_scatters_loops(data, ofl, h['duration'], lor, table, S, scale, h['tof_range'],
                                              mash, shift)
    assert kept.astype('<u4').tobytes() == want_bytes
    assert 0 < kept.shape[0] < ev.shape[0]


def test_file_drop_ins_round_trip(oracle, tmp_path):
    """ExportPrompts and InjectScatters on files: the injected datafile and its
    header (event count, scatter flag, new names) as the reference writes them."""
    lor, sino = _scanner(64, 2)
    h = castor.listmode_header('scan_df.Cdf', 20000, duration=100.0, tof_range=600.0, random=True)
    rng = np.random.default_rng(3)
    ev = castor.synthetic_events(h, 20000, 64, NR, rng)
    castor.write_listmode(tmp_path, 'scan_df.Cdh', h, ev)
    P, R, Nrm = castor.ExportPrompts(str(tmp_path), 'scan_df.Cdh', lor, sino, 3, mash=2)
    assert P.shape == (lor[:, :, 1].max() + 1, lor[:, :, 0].max() + 1, NR * NR, 3) and Nrm is None
    want = oracle.export_prompts_loops(ev.astype('<u4').tobytes(), (0, 0, 1, 0, 1, 1), 100.0, lor, sino, 3,
                                       600.0, 2, 0)
    assert np.array_equal(P, want[0]) and np.array_equal(R, want[1])
    scale = np.array([1.5, 0.75, 1.25], np.float32)
    np.save(tmp_path / 'ScaleFactors.npy', scale)
    S = rng.uniform(0.1, 3.0, size=(3,) + P.shape[:3]).astype(np.float32)
    for b in range(3):
        np.savez_compressed(tmp_path / f'SSS_mashed_bin{b}.npz', S[b])
    castor.InjectScatters(str(tmp_path), 'scan_df.Cdh', lor, sino, str(tmp_path), None, False, mash=2)
    nh = castor.ReadHeaderFile(tmp_path / 'scan_scatter_df.Cdh')
    assert nh['scatter_flag'] and nh['data_filename'] == 'scan_scatter_df.Cdf'
    got = (tmp_path / 'scan_scatter_df.Cdf').read_bytes()
    want_bytes = oracle.inject_scatters_loops(ev.astype('<u4').tobytes(), (0, 0, 1, 0, 1, 1), 100.0, lor, sino, S,
                                              scale, 600.0, 2, 0)
    assert got == want_bytes
    assert nh['nb_events'] == len(got) // (4 * castor.event_words(nh))


def test_out_of_table_crystal_is_an_error():
    lor, sino = _scanner(64, 1)
    h = castor.listmode_header('x_df.Cdf', 0)
    ev = castor.synthetic_events(h, 100, 64, NR, np.random.default_rng(1))
    ev[50, 3] = 64 * NR + 5          # a ring the tables do not have
    from opensss_amd._native import SssError
    with pytest.raises(SssError, match='outside the tables'):
        castor.histogram_events(ev, h, lor, sino, 4, 1000.0)
