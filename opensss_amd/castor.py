"""CASToR list-mode datafiles on either side of the scatter path, on the GPU.

Drop-ins for ``Python/Timber``: ``ReadHeaderFile`` / ``WriteHeaderFile``
(``Headerfile.py:4-143``), ``GenerateEventClass`` / ``GenerateListEventInfo``
/ ``ParseCASToRIDs`` (``Coincidence.py:21-159``), ``ExportPrompts``
(``Datafile.py:29-115``: events -> prompts, randoms and normalisation
sinograms) and ``InjectScatters`` (``Datafile.py:117-240``: events ->
a new datafile with each kept event's scatter rate).  The per-event work runs
in ``sss_listmode_histogram`` / ``sss_listmode_inject``; here are the header
text, the event layout and the file handling.

Reference behaviours kept: a header flag line counts as set whatever its
digit (``bool('0')`` is true); the randoms sinogram takes word 0 of the
cell's last event (the coincidence time) times the duration over the bins and
the normalisation word 1 of that event; prompts are uint16 counts (wrapping);
the scale factors of the injection are divided by the bin width and the rate
factor formed in float32 (numpy's promotion of a float32 scalar with Python
floats); ``InjectScatters`` needs a datafile without scatter corrections
(the reference's struct packing fails on one that has them) and both need
TOF (the reference reads the TOF word unconditionally).
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import _native


# ---- header (Headerfile.py) ----

def _find(pattern, text):
    return re.findall(pattern, text)


def ReadHeaderFile(HeaderFilePath: str) -> dict:
    """Header dictionary with the reference's keys (Headerfile.py:4-95)."""
    with open(HeaderFilePath, 'r') as fh:
        header = fh.read()

    def flag(name):
        v = _find(name + r': (\d)', header)
        return bool(v[0]) if len(v) != 0 else False

    max_lines = _find(r'Maximum number of lines per event: (\d*)', header)
    out = {
        'data_filename': _find(r'Data filename: (\S*)', header)[0],
        'nb_events': int(_find(r'Number of events: (\d*)', header)[0]),
        'data_mode': _find(r'Data mode: (\w*[-]?\w*)', header)[0],
        'data_type': _find(r'Data type: (\w*)', header)[0],
        'start_time': float(_find(r'Start time \(s\): (\d*[.]?\d*)', header)[0]),
        'duration': float(_find(r'Duration \(s\): (\d*[.]?\d*)', header)[0]),
        'scanner_name': _find(r'Scanner name: (\S*)', header)[0],
        'isotope': _find(r'Isotope: (\S*)', header)[0],
        'calibration_factor': float(_find(r'Calibration factor: (\d*[.]?\d*)', header)[0]),
        'nb_max_lines': int(max_lines[0]) if len(max_lines) != 0 else 1,
        'attenuation_flag': flag('Attenuation correction flag'),
        'normalization_flag': flag('Normalization correction flag'),
        'scatter_flag': flag('Scatter correction flag'),
        'random_flag': flag('Random correction flag'),
        'tof_flag': flag('TOF information flag'),
        'tof_resolution': None, 'tof_bin_size': None, 'tof_nb_bins': None, 'tof_range': None,
    }
    if out['tof_flag']:
        out['tof_resolution'] = float(_find(r'TOF resolution \(ps\): (\d*[.]?\d*)', header)[0])
        if 'histo' in out['data_mode']:
            out['tof_nb_bins'] = int(_find(r'Histo TOF number of bins: (\d*)', header)[0])
            out['tof_bin_size'] = float(_find(r'Histo TOF bin size \(ps\): (\d*[.]?\d*)', header)[0])
        elif 'list' in out['data_mode']:
            out['tof_range'] = float(_find(r'List TOF measurement range \(ps\): (\d*[.]?\d*)', header)[0])
    return out


def WriteHeaderFile(Header: dict, HeaderFilePath: str) -> None:
    """The header text of Headerfile.py:97-143."""
    lines = [f'Data filename: {Header["data_filename"]}',
             f'Number of events: {Header["nb_events"]}',
             f'Data mode: {Header["data_mode"]}',
             f'Data type: {Header["data_type"]}',
             f'Start time (s): {Header["start_time"]}',
             f'Duration (s): {Header["duration"]}',
             f'Scanner name: {Header["scanner_name"]}',
             f'Isotope: {Header["isotope"]}',
             f'Calibration factor: {Header["calibration_factor"]}']
    if Header['nb_max_lines'] > 1:
        lines.append(f'Maximum number of lines per event: {Header["nb_max_lines"]}')
    for key, text in (('attenuation_flag', 'Attenuation'), ('normalization_flag', 'Normalization'),
                      ('scatter_flag', 'Scatter'), ('random_flag', 'Random')):
        if Header[key]:
            lines.append(f'{text} correction flag: 1')
    if Header['tof_flag']:
        lines.append('TOF information flag: 1')
        lines.append(f'TOF resolution (ps): {Header["tof_resolution"]}')
        if 'histo' in Header['data_mode']:
            lines.append(f'Histo TOF number of bins: {Header["tof_nb_bins"]}')
            lines.append(f'Histo TOF bin size (ps): {Header["tof_bin_size"]}')
        elif 'list' in Header['data_mode']:
            lines.append(f'List TOF measurement range (ps): {Header["tof_range"]}')
    with open(HeaderFilePath, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')


# ---- event layout (Coincidence.py) ----

def GenerateListEventInfo(AttenuationFlag, ScatterFlag, RandomFlag, NormalizationFlag, TOFFlag, MaxNbLines):
    """(struct format, positions (scatter, random, norm, TOF, first crystal ID))
    of a list-mode event, Coincidence.py:45-95; every field is 4 bytes."""
    nb_lines_flag = MaxNbLines > 1
    fields = '<1L'
    nc = AttenuationFlag + ScatterFlag + RandomFlag + NormalizationFlag + TOFFlag
    if nc > 0:
        fields += f'{nc}f'
    if nb_lines_flag:
        fields += '1I'
    fields += f'{2 * MaxNbLines}L'
    a = int(AttenuationFlag)
    pos = (1 + a if ScatterFlag else None,
           1 + a + ScatterFlag if RandomFlag else None,
           1 + a + ScatterFlag + RandomFlag if NormalizationFlag else None,
           1 + a + ScatterFlag + RandomFlag + NormalizationFlag if TOFFlag else None,
           1 + a + ScatterFlag + RandomFlag + NormalizationFlag + TOFFlag + nb_lines_flag)
    return fields, pos


def GenerateEventClass(DataHeader: dict):
    if 'list' in DataHeader['data_mode']:
        return GenerateListEventInfo(DataHeader['attenuation_flag'], DataHeader['scatter_flag'],
                                     DataHeader['random_flag'], DataHeader['normalization_flag'],
                                     DataHeader['tof_flag'], DataHeader['nb_max_lines'])
    raise NotImplementedError('Handling of histogram data is not yet implemented')


def event_words(header: dict) -> int:
    """4-byte words per event record."""
    nc = (header['attenuation_flag'] + header['scatter_flag'] + header['random_flag'] +
          header['normalization_flag'] + header['tof_flag'])
    return 1 + int(nc) + int(header['nb_max_lines'] > 1) + 2 * int(header['nb_max_lines'])


def ParseCASToRIDs(event, field_pos, NrCrystalsPerRing, Shift=0):
    """(detector1, ring1, detector2, ring2, TOF) of one unpacked event
    (Coincidence.py:97-159): detectors from the IDs without the shift, the
    ring order and TOF sign from the shifted transaxial indices."""
    c0, c1 = event[field_pos[4]], event[field_pos[4] + 1]
    tof_info = event[field_pos[3]]
    id1, id2 = (c0, c1) if c0 < c1 else (c1, c0)
    tof_tmp = tof_info if c0 < c1 else -tof_info
    t1, t2 = id1 % NrCrystalsPerRing, id2 % NrCrystalsPerRing
    d1, d2 = (t1, t2) if t1 < t2 else (t2, t1)
    s1, s2 = id1 % NrCrystalsPerRing - Shift, id2 % NrCrystalsPerRing - Shift
    if s1 < 0:
        s1 += NrCrystalsPerRing
    if s2 < 0:
        s2 += NrCrystalsPerRing
    if s1 < s2:
        return d1, id1 // NrCrystalsPerRing, d2, id2 // NrCrystalsPerRing, tof_tmp
    return d1, id2 // NrCrystalsPerRing, d2, id1 // NrCrystalsPerRing, -tof_tmp


class _Layout(_native.ctypes.Structure):
    _fields_ = [('words', _native.ctypes.c_int32), ('tof_word', _native.ctypes.c_int32),
                ('id_word', _native.ctypes.c_int32), ('scatter_word', _native.ctypes.c_int32)]


def _layout(header: dict, scatter_word: int = 0):
    if not header['tof_flag']:
        raise ValueError('the list-mode datafile has no TOF information (the reference reads the TOF word of '
                         'every event)')
    _, pos = GenerateEventClass(header)
    return _Layout(event_words(header), pos[3], pos[4], scatter_word)


def read_events(DataPath: str, header: dict) -> np.ndarray:
    """The datafile as (nb_events, words) uint32, after the reference's size check."""
    path = os.path.join(DataPath, header['data_filename'])
    words = event_words(header)
    size = os.path.getsize(path)
    if size / (4 * words) != header['nb_events']:
        raise ValueError('The number of events in the file is not consistent with the header')
    return np.fromfile(path, dtype='<u4').reshape(-1, words)


def _tables(LORCoordinates, RingIndex):
    lor = np.ascontiguousarray(np.asarray(LORCoordinates, dtype=np.int16))
    ri = np.ascontiguousarray(np.asarray(RingIndex, dtype=np.int16))
    if lor.ndim != 3 or lor.shape[0] != lor.shape[1] or lor.shape[2] != 2 or ri.ndim != 2 or ri.shape[0] != ri.shape[1]:
        raise ValueError('LORCoordinates must be (N, N, 2) and the ring table square')
    return lor, ri


def histogram_events(events, header, LORCoordinates, RingIndex, TOF_bins, TOF_range, mash=1, Shift=0, *,
                     context=None):
    """ExportPrompts' binning of an (n, words) uint32 event array: (prompts
    uint16 (rad, ang, planes, TOF_bins), randoms f32 or None, norm f32 or None)."""
    lor, ri = _tables(LORCoordinates, RingIndex)
    ev = np.ascontiguousarray(events, dtype=np.uint32)
    lay = _layout(header)
    n_rad = int(lor[:, :, 1].max()) + 1
    n_ang = int(lor[:, :, 0].max()) + 1
    n_planes = int(ri.max()) + 1
    prompts = np.zeros((n_rad, n_ang, n_planes, int(TOF_bins)), dtype=np.uint16)
    randoms = np.zeros((n_rad, n_ang, n_planes), np.single) if header['random_flag'] else None
    norm = np.ones((n_rad, n_ang, n_planes), np.single) if header['normalization_flag'] else None
    flags = (_native.SSS_LM_RANDOMS if randoms is not None else 0) | (_native.SSS_LM_NORM if norm is not None else 0)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    _native.check(lib.sss_listmode_histogram(
        ctx.handle, _native.ptr(ev), ev.shape[0], _native.ctypes.byref(lay), lor.shape[0] * int(mash), int(Shift),
        int(mash), lor.shape[0], _native.ptr(lor), ri.shape[0], _native.ptr(ri), n_rad, n_ang, n_planes,
        int(TOF_bins), float(TOF_range), float(header['duration']), flags, _native.ptr(prompts),
        None if randoms is None else _native.ptr(randoms), None if norm is None else _native.ptr(norm)),
        'sss_listmode_histogram')
    return prompts, randoms, norm


def ExportPrompts(DataPath: str, DataHeader: str, LORCoordinates, SinogramIndex, TOF_bins: int,
                  TOF_range: float = None, mash: int = 1, Shift: int = 0, *, context=None):
    """Prompts (uint16), randoms and normalisation sinograms of a CASToR
    list-mode datafile (Datafile.py:29-115)."""
    header = ReadHeaderFile(os.path.join(DataPath, DataHeader))
    events = read_events(DataPath, header)
    if TOF_range is None:
        TOF_range = header['tof_range']
    return histogram_events(events, header, LORCoordinates, SinogramIndex, TOF_bins, TOF_range, mash, Shift,
                            context=context)


SPEED_OF_LIGHT_IN_MM_PER_PS = 0.299792458


def injection_factors(Scale, TOF_range, scan_duration):
    """Per-bin float32 rate factors, Datafile.py:161,193,222: Scale /= the bin
    width (in place on the float32 array), then Scale[b] * 1/(c/2) / duration
    with numpy's scalar promotion."""
    scale = np.array(Scale, dtype=np.float32, copy=True)
    scale /= (2 * TOF_range / scale.shape[0])
    scaling_factor = 1. / (SPEED_OF_LIGHT_IN_MM_PER_PS / 2)
    return np.array([scale[b] * scaling_factor / scan_duration for b in range(scale.shape[0])], dtype=np.float32)


def inject_events(events, header, LORCoordinates, RingIndex, scatter_sinograms, factors, TOF_range, mash=1,
                  Shift=0, *, context=None):
    """InjectScatters' selection of an (n, words) uint32 event array: the kept
    events as (m, words + 1) uint32 with the scatter rate inserted."""
    if header['scatter_flag']:
        raise ValueError('the datafile already holds scatter corrections')
    lor, ri = _tables(LORCoordinates, RingIndex)
    ev = np.ascontiguousarray(events, dtype=np.uint32)
    new_header = dict(header, scatter_flag=True)
    _, new_pos = GenerateEventClass(new_header)
    lay = _layout(header, scatter_word=new_pos[0])
    S = np.ascontiguousarray(np.asarray(scatter_sinograms, dtype=np.float32))
    if S.ndim != 4 or S.shape[0] != len(factors):
        raise ValueError('scatter sinograms must be (TOF bins, rad, ang, planes)')
    f = np.ascontiguousarray(np.asarray(factors, dtype=np.float32))
    out = np.empty((ev.shape[0], lay.words + 1), dtype=np.uint32)
    n_out = _native.ctypes.c_int64(0)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    _native.check(lib.sss_listmode_inject(
        ctx.handle, _native.ptr(ev), ev.shape[0], _native.ctypes.byref(lay), lor.shape[0] * int(mash), int(Shift),
        int(mash), lor.shape[0], _native.ptr(lor), ri.shape[0], _native.ptr(ri), S.shape[1], S.shape[2], S.shape[3],
        S.shape[0], float(TOF_range), _native.ptr(f), _native.ptr(S), 0, _native.ptr(out),
        _native.ctypes.byref(n_out)), 'sss_listmode_inject')
    return out[:n_out.value]


def InjectScatters(DataPath: str, DataHeader: str, LORCoordinates, SinogramIndex, ScatterPath: str, LookUpTable,
                   SpanFlag: bool, TOF_range: float = None, NewPath: str = None, mash: int = 1, Shift: int = 0, *,
                   context=None):
    """Writes ``<name>scatter_<ext>`` datafile and header with every kept
    event's scatter rate (Datafile.py:117-240)."""
    header = ReadHeaderFile(os.path.join(DataPath, DataHeader))
    scale = np.load(os.path.join(ScatterPath, 'ScaleFactors.npy'))
    n_bins = scale.shape[0]
    events = read_events(DataPath, header)
    new_header = header.copy()
    new_header['scatter_flag'] = True
    original = header['data_filename']
    new_header['data_filename'] = original[:-6] + 'scatter_' + original[-6:]
    new_header_name = DataHeader[:-6] + 'scatter_' + DataHeader[-6:]
    if TOF_range is None:
        TOF_range = header['tof_range']
    else:
        new_header['tof_range'] = TOF_range
    if NewPath is None:
        NewPath = DataPath
    factors = injection_factors(scale, TOF_range, header['duration'])
    SI = np.asarray(SinogramIndex)
    lor = np.asarray(LORCoordinates)
    planes = SI.shape[0] if SpanFlag else int(np.max(SI)) + 1
    size = (n_bins, int(np.max(lor[:, :, 1])) + 1, int(np.max(lor[:, :, 0])) + 1, planes)
    S = np.zeros(size, dtype=np.float32)
    for i in range(n_bins):
        with np.load(os.path.join(ScatterPath, f'SSS_mashed_bin{i}.npz')) as z:
            S[i] = z['arr_0']
    ring_table = SI if SI.shape[0] == SI.shape[1] else np.asarray(LookUpTable)
    kept = inject_events(events, header, LORCoordinates, ring_table, S, factors, TOF_range, mash, Shift,
                         context=context)
    kept.astype('<u4').tofile(os.path.join(NewPath, new_header['data_filename']))
    new_header['nb_events'] = int(kept.shape[0])
    WriteHeaderFile(new_header, os.path.join(NewPath, new_header_name))


# ---- synthetic list-mode data (tests, benchmarks) ----

def listmode_header(data_filename, nb_events, scanner_name='synthetic', duration=600.0, tof_range=1000.0,
                    attenuation=False, scatter=False, random=False, normalization=False, max_lines=1,
                    tof_resolution=214.0):
    """A list-mode header dictionary with TOF."""
    return {'data_filename': data_filename, 'nb_events': int(nb_events), 'data_mode': 'list-mode',
            'data_type': 'PET', 'start_time': 0.0, 'duration': float(duration), 'scanner_name': scanner_name,
            'isotope': 'F-18', 'calibration_factor': 1.0, 'nb_max_lines': int(max_lines),
            'attenuation_flag': bool(attenuation), 'normalization_flag': bool(normalization),
            'scatter_flag': bool(scatter), 'random_flag': bool(random), 'tof_flag': True,
            'tof_resolution': float(tof_resolution), 'tof_bin_size': None, 'tof_nb_bins': None,
            'tof_range': float(tof_range)}


def synthetic_events(header, n_events, n_crystals_per_ring, n_rings, rng, tof_spread=1.2):
    """(n_events, words) uint32 events: random crystal pairs over the rings,
    TOF uniform over +-tof_spread times the range (float32), times and
    corrections random."""
    words = event_words(header)
    _, pos = GenerateEventClass(header)
    ev = np.zeros((n_events, words), dtype=np.uint32)
    ev[:, 0] = rng.integers(0, 2 ** 20, size=n_events, dtype=np.uint32)
    n_corr = pos[4] - 1 - int(header['nb_max_lines'] > 1)
    if n_corr > 0:
        ev[:, 1:1 + n_corr] = rng.uniform(0.5, 2.0, size=(n_events, n_corr)).astype(np.float32).view(np.uint32)
    if header['nb_max_lines'] > 1:
        ev[:, pos[4] - 1] = 1
    n_ids = n_crystals_per_ring * n_rings
    ev[:, pos[4]:] = rng.integers(0, n_ids, size=(n_events, words - pos[4]), dtype=np.uint32)
    tof = rng.uniform(-tof_spread, tof_spread, size=n_events) * header['tof_range']
    ev[:, pos[3]] = tof.astype(np.float32).view(np.uint32)
    return ev


def write_listmode(path, header_name, header, events):
    """Writes the header and the events (little-endian words) next to it."""
    WriteHeaderFile(header, os.path.join(path, header_name))
    np.ascontiguousarray(events, dtype='<u4').tofile(os.path.join(path, header['data_filename']))
