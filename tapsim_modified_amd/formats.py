"""TAPSim's on-disk inputs for the hot path: the species configuration
(`.aly`, ASCII) and the node file (coordinates, ids, optional unique numbers
and potentials; ASCII or BINARY), read and written as the reference's
File_Io does (tapsim/file_io.cpp), so that unmodified TAPSim inputs drive
this library and its outputs can be diffed with the reference's.

    temperature, species = read_config("example.aly")     # {id: dict(...)}
    table = tapsim_modified_amd.species_table(species)      # -> Grid3D(...)
    nodes = read_node_file("tip.nodes")                     # xyz, ids, numbers, phi
"""
import struct

import numpy as np

# File_Io's configuration keys and units (file_io.cpp:38-53)
CONFIG_KEYS = (
    ("TEMPERATURE", "[K]"),
    ("ID", "[1]"),
    ("NAME", "[1]"),
    ("CHARGE_DENSITY", "[C/m³]"),
    ("DIELECTRICITY", "[1]"),
    ("REMOVABLE", "[1]"),
    ("NEUMANN_BOUNDARY", "[1]"),
    ("DIRICHLET_BOUNDARY", "[1]"),
    ("POTENTIAL", "[V]"),
    ("MASS", "[u]"),
    ("EVAPORATION_CHARGE_STATE", "[1]"),
    ("EVAPORATION_FIELD_STRENGTH", "[V/m]"),
    ("EVAPORATION_ACTIVATION_ENERGY", "[eV]"),
)

# key -> (field of species_table(), parser); name is kept under "name"
_FIELDS = {
    "ID": ("id", int),
    "NAME": ("name", str),
    "CHARGE_DENSITY": ("charge_density", float),
    "DIELECTRICITY": ("epsilon", float),
    "REMOVABLE": ("removable", int),
    "NEUMANN_BOUNDARY": ("neumann", int),
    "DIRICHLET_BOUNDARY": ("dirichlet", int),
    "POTENTIAL": ("phi", float),
    "MASS": ("mass", float),
    "EVAPORATION_CHARGE_STATE": ("evap_charge", int),
    "EVAPORATION_FIELD_STRENGTH": ("evap_field", float),
    "EVAPORATION_ACTIVATION_ENERGY": ("evap_energy", float),
}

# Configuration::NodeNumber's invalid value (configuration.h:82-87: numbers
# start at 1, 0 is invalid)
INVALID_NUMBER = 0


class FormatError(RuntimeError):
    """The input does not follow the format (the reference throws
    std::runtime_error("File_Io::...") at the same places)."""


def _f32(x):
    return float(np.float32(x))


def _parse(kind, text, where):
    try:
        if kind is int:
            return int(text.split()[0])
        if kind is float:
            return _f32(float(text.split()[0]))
        return text
    except (ValueError, IndexError):
        raise FormatError("%s: cannot read %r" % (where, text))


def read_config(path):
    """File_Io::readConfig (file_io.cpp:212-463), ASCII files: returns
    (temperature, {id: fields}) with the fields species_table() takes plus
    "name".  Blocks are separated by empty lines, lines starting with '#'
    are skipped, keys and values split at " = "; a block of one key is the
    temperature, a block of twelve a species."""
    with open(path, "r", encoding="utf-8", errors="surrogateescape") as f:
        lines = f.read().split("\n")
    pos = 0
    while pos < len(lines) and (not lines[pos] or lines[pos][0] == "#"):
        pos += 1
    if pos >= len(lines):
        raise FormatError("File_Io::readConfig(): no header")
    if lines[pos] == "BINARY":
        raise FormatError("File_Io::readConfig(): BINARY configuration files are not read here; write them as ASCII")
    if lines[pos] != "ASCII":
        raise FormatError("File_Io::readConfig(): header must be ASCII")
    pos += 1
    temperature = 0.0
    species = {}
    n = len(lines)
    while pos < n:
        while pos < n and (not lines[pos] or lines[pos][0] == "#"):
            pos += 1
        block = {}
        while pos < n and lines[pos]:
            line = lines[pos]
            if line[0] != "#":
                cut = line.find(" = ")
                if cut < 0:
                    raise FormatError("File_Io::readConfig(): no ' = ' in %r" % line)
                key, value = line[:cut].strip(), line[cut + 3:].strip()
                if key in block:
                    raise FormatError("File_Io::readConfig(): duplicate key %s" % key)
                block[key] = value
            pos += 1
        if not block:
            continue
        if len(block) == 1:
            key, value = next(iter(block.items()))
            if key != CONFIG_KEYS[0][0]:
                raise FormatError("File_Io::readConfig(): a one-line block must be the temperature")
            temperature = _parse(float, value, key)
        elif len(block) == 12:
            entry = {}
            for key, value in block.items():
                if key not in _FIELDS:
                    raise FormatError("File_Io::readConfig(): unknown key %s" % key)
                field, kind = _FIELDS[key]
                entry[field] = _parse(kind, value, key)
            sid = entry.pop("id")
            if not 0 <= sid <= 127:
                raise FormatError("File_Io::readConfig(): id %d outside 0..127" % sid)
            species[sid] = entry
        else:
            raise FormatError("File_Io::readConfig(): a block must hold 1 or 12 keys")
    return temperature, species


def _g8(x):
    """a float as std::ostream with precision(8) writes it (general format)"""
    return "%.8g" % _f32(x)


def write_config(path, temperature, species):
    """File_Io::writeConfig, ASCII mode (file_io.cpp:116-174): the header
    comments, then the temperature and every species in ascending id."""
    out = ["# global parameters", "#", "# %s %s" % CONFIG_KEYS[0], "#", "# species dependant parameters", "#"]
    out += ["# %s %s" % CONFIG_KEYS[i] for i in range(1, 13)]
    out += ["ASCII", "", "%s = %s" % (CONFIG_KEYS[0][0], _g8(temperature)), ""]
    for sid in sorted(species):
        s = species[sid]
        out += [
            "ID = %d" % sid,
            "NAME = %s" % s.get("name", ""),
            "CHARGE_DENSITY = %s" % _g8(s.get("charge_density", 0.0)),
            "DIELECTRICITY = %s" % _g8(s.get("epsilon", 1.0)),
            "REMOVABLE = %d" % int(bool(s.get("removable", 0))),
            "NEUMANN_BOUNDARY = %d" % int(bool(s.get("neumann", 0))),
            "DIRICHLET_BOUNDARY = %d" % int(bool(s.get("dirichlet", 0))),
            "POTENTIAL = %s" % _g8(s.get("phi", 0.0)),
            "MASS = %s" % _g8(s.get("mass", 0.0)),
            "EVAPORATION_CHARGE_STATE = %d" % int(s.get("evap_charge", 0)),
            "EVAPORATION_FIELD_STRENGTH = %s" % _g8(s.get("evap_field", 0.0)),
            "EVAPORATION_ACTIVATION_ENERGY = %s" % _g8(s.get("evap_energy", 0.0)),
            "",
        ]
    with open(path, "w", encoding="utf-8", newline="\n") as f:
        f.write("\n".join(out) + "\n")


def _renumber(numbers):
    """readGeomGrid's renumbering (file_io.cpp:975-1004): valid numbers must
    be unique and stay; invalid ones get the next numbers after the largest
    valid one (from 1), in node order."""
    valid = numbers != INVALID_NUMBER
    if len(np.unique(numbers[valid])) != int(valid.sum()):
        raise FormatError("File_Io::readGeomGrid(): non unique numbering detected")
    nxt = int(numbers[valid].max()) + 1 if valid.any() else 1
    out = numbers.copy()
    idx = np.nonzero(~valid)[0]
    out[idx] = (np.arange(nxt, nxt + len(idx), dtype=np.uint64) & 0xFFFFFFFF).astype(np.uint32)
    return out


def read_node_file(path):
    """File_Io::readGeomGrid (file_io.cpp:822-1011): the header "ASCII n
    numbers potentials" or "BINARY ..." (lines starting with '#' before it
    are skipped), then per node x y z id [number] [phi] (tab separated in
    ASCII; float32 x3, int16, [uint32], [float32] packed in BINARY).
    Returns dict(xyz float32 (n,3), ids int16, numbers uint32 (renumbered
    as the reference does), phi float32 (0 when absent), with_numbers,
    with_potentials); with_potentials repeats with_numbers, as the
    reference reports it (file_io.cpp:1008)."""
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0
    while True:
        end = raw.find(b"\n", pos)
        if end < 0:
            raise FormatError("File_Io::readGeomGrid(): no header")
        line = raw[pos:end].decode("latin-1")
        pos = end + 1
        if line and line[0] != "#":
            break
    if line.startswith("ASCII"):
        ascii_mode, rest = True, line[5:]
    elif line.startswith("BINARY"):
        ascii_mode, rest = False, line[6:]
    else:
        raise FormatError("File_Io::readGeomGrid(): header must start with ASCII or BINARY")
    try:
        head = rest.split()
        n, with_numbers, with_potentials = int(head[0]), int(head[1]) != 0, int(head[2]) != 0
    except (ValueError, IndexError):
        raise FormatError("File_Io::readGeomGrid(): bad header %r" % line)
    xyz = np.zeros((n, 3), np.float32)
    ids = np.full(n, -1, np.int16)
    numbers = np.full(n, INVALID_NUMBER, np.uint32)
    phi = np.zeros(n, np.float32)
    if ascii_mode:
        lines = raw[pos:].decode("latin-1").split("\n")
        li = 0
        for i in range(n):
            while li < len(lines) and lines[li][:1] == "#":
                li += 1
            if li >= len(lines):
                raise FormatError("File_Io::readGeomGrid(): file ends at node %d" % i)
            vals = lines[li].split("\t")
            li += 1
            need = 4 + with_numbers + with_potentials
            if len(vals) < need:
                raise FormatError("File_Io::readGeomGrid(): node %d has %d values" % (i, len(vals)))
            xyz[i] = [_parse(float, v, "coords") for v in vals[:3]]
            ids[i] = _parse(int, vals[3], "id")
            k = 4
            if with_numbers:
                numbers[i] = _parse(int, vals[k], "number")
                k += 1
            if with_potentials:
                phi[i] = _parse(float, vals[k], "potential")
    else:
        fmt = "<3fh" + ("I" if with_numbers else "") + ("f" if with_potentials else "")
        rec = struct.calcsize(fmt)
        if len(raw) - pos < rec * n:
            raise FormatError("File_Io::readGeomGrid(): binary body too short")
        dt = [("xyz", "<f4", 3), ("id", "<i2")]
        if with_numbers:
            dt.append(("number", "<u4"))
        if with_potentials:
            dt.append(("phi", "<f4"))
        body = np.frombuffer(raw, dtype=np.dtype(dt), count=n, offset=pos)
        xyz[:] = body["xyz"]
        ids[:] = body["id"]
        if with_numbers:
            numbers[:] = body["number"]
        if with_potentials:
            phi[:] = body["phi"]
    return dict(xyz=xyz, ids=ids, numbers=_renumber(numbers), phi=phi, with_numbers=with_numbers,
                with_potentials=with_numbers)


def _e8(x):
    """a float as std::ostream with scientific, showpoint and precision(8)"""
    return "%.8e" % _f32(x)


def write_node_file(path, xyz, ids, numbers=None, phi=None, mode="ASCII"):
    """File_Io::writeGeomGrid (file_io.cpp:765-820): ASCII (x y z id [number]
    [phi], tab separated, floats in scientific notation with 8 digits) or
    BINARY (the same fields packed)."""
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    ids = np.asarray(ids, np.int16)
    n = len(xyz)
    wn, wp = int(numbers is not None), int(phi is not None)
    head = ("%s %d %d %d\n" % (mode, n, wn, wp)).encode("latin-1")
    if mode == "ASCII":
        rows = []
        for i in range(n):
            parts = [_e8(xyz[i, 0]), _e8(xyz[i, 1]), _e8(xyz[i, 2]), "%d" % ids[i]]
            if wn:
                parts.append("%d" % int(numbers[i]))
            if wp:
                parts.append(_e8(phi[i]))
            rows.append("\t".join(parts) + "\n")
        body = "".join(rows).encode("latin-1")
    elif mode == "BINARY":
        dt = [("xyz", "<f4", 3), ("id", "<i2")]
        if wn:
            dt.append(("number", "<u4"))
        if wp:
            dt.append(("phi", "<f4"))
        rec = np.zeros(n, np.dtype(dt))
        rec["xyz"] = xyz
        rec["id"] = ids
        if wn:
            rec["number"] = np.asarray(numbers, np.uint32)
        if wp:
            rec["phi"] = np.asarray(phi, np.float32)
        body = rec.tobytes()
    else:
        raise FormatError("File_Io::writeGeomGrid(): mode must be ASCII or BINARY")
    with open(path, "wb") as f:
        f.write(head + body)


# ---------------------------------------------------------------- results stream

RESULTS_VERSION = "1.0"  # File_Io::LogResults::_fileVersion (logResults.cpp:34)

# one record per event, in the column order of the results file
RESULT_FIELDS = [
    ("index", "<u4"), ("id", "<i2"), ("number", "<u4"), ("voltage", "<f4"), ("start", "<f4", 3), ("stop", "<f4", 3),
    ("tof", "<f4"), ("probability", "<f4"), ("pot_before", "<f4"), ("field_before", "<f4", 3), ("pot_after", "<f4"),
    ("field_after", "<f4", 3), ("normal", "<f4", 3), ("apex", "<f4", 3),
]
RESULT_DTYPE = np.dtype(RESULT_FIELDS)

_RESULT_COLUMNS = (
    ("index", "[1]", "int"), ("id", "[1]", "int"), ("number", "[1]", "unsigned int"), ("voltage", "[V]", "float"),
    ("startX", "[m]", "float"), ("startY", "[m]", "float"), ("startZ", "[m]", "float"),
    ("stopX", "[m]", "float"), ("stopY", "[m]", "float"), ("stopZ", "[m]", "float"), ("tof", "[s]", "float"),
    ("probability", "[1]", "float"), ("potentialBefore", "[V]", "float"), ("fieldBeforeX", "[V/m]", "float"),
    ("fieldBeforeY", "[V/m]", "float"), ("fieldBeforeZ", "[V/m]", "float"), ("potentialAfter", "[V]", "float"),
    ("fieldAfterX", "[V/m]", "float"), ("fieldAfterY", "[V/m]", "float"), ("fieldAfterZ", "[V/m]", "float"),
    ("normalX", "[m]", "float"), ("normalY", "[m]", "float"), ("normalZ", "[m]", "float"),
    ("apexX", "[m]", "float"), ("apexY", "[m]", "float"), ("apexZ", "[m]", "float"),
)


def result_records(events, numbers=None, first_index=1):
    """Records of the results file from tsg_evaporate's events: the event
    index counts from first_index (Process::evaporation's ++eventCnt), the
    unique number is numbers[node] (a node file's numbers) or node + 1."""
    rec = np.zeros(len(events), RESULT_DTYPE)
    for k, e in enumerate(events):
        r = rec[k]
        r["index"] = first_index + k
        r["id"] = e.id
        r["number"] = int(numbers[e.node]) if numbers is not None else e.node + 1
        r["voltage"] = e.voltage
        r["start"] = list(e.start)
        r["stop"] = list(e.stop)
        r["tof"] = e.tof
        r["probability"] = e.probability
        r["pot_before"] = e.pot_before
        r["field_before"] = list(e.field_before)
        r["pot_after"] = e.pot_after
        r["field_after"] = list(e.field_after)
        r["normal"] = list(e.normal)
        r["apex"] = list(e.apex)
    return rec


def write_results(path, records, mode="ASCII", header=""):
    """File_Io::LogResults (logResults.cpp:135-366), one file: "TAPSIM
    RESULTS DATA", the version, an optional header, the column comments and
    the mode line, then one line per event (tab separated, floats in
    scientific notation with six digits) or one packed record per event."""
    records = np.asarray(records, RESULT_DTYPE)
    head = ["TAPSIM RESULTS DATA", "VERSION " + RESULTS_VERSION]
    text = "\n".join(head) + "\n" + header
    if mode == "BINARY":
        text += "".join("# %s %s (%s)\n" % c for c in _RESULT_COLUMNS[:12])
        # the reference labels both potentials "[1]" in binary files
        text += "# potentialBefore [1] (float)\n"
        text += "".join("# %s %s (%s)\n" % c for c in _RESULT_COLUMNS[13:16])
        text += "# potentialAfter [1] (float)\n"
        text += "".join("# %s %s (%s)\n" % c for c in _RESULT_COLUMNS[17:])
        text += "BINARY\n"
        body = records.tobytes()
    elif mode == "ASCII":
        text += "".join("# %s %s\n" % c[:2] for c in _RESULT_COLUMNS)
        text += "ASCII\n"
        rows = []
        for r in records:
            vals = ["%d" % r["index"], "%d" % r["id"], "%d" % r["number"], "%e" % float(r["voltage"])]
            vals += ["%e" % float(v) for v in r["start"]] + ["%e" % float(v) for v in r["stop"]]
            vals += ["%e" % float(r[k]) for k in ("tof", "probability", "pot_before")]
            vals += ["%e" % float(v) for v in r["field_before"]] + ["%e" % float(r["pot_after"])]
            vals += ["%e" % float(v) for v in r["field_after"]] + ["%e" % float(v) for v in r["normal"]]
            vals += ["%e" % float(v) for v in r["apex"]]
            rows.append("\t".join(vals) + "\n")
        body = "".join(rows).encode("latin-1")
    else:
        raise FormatError("File_Io::LogResults::header(): Invalid output mode!")
    with open(path, "wb") as f:
        f.write(text.encode("latin-1") + body)


# ---------------------------------------------------------------- trajectory stream

TRAJECTORY_VERSION = "1.0"
_TRAJ_COLUMNS = (("time", "[s]", "float"), ("positionX", "[m]", "float"), ("positionY", "[m]", "float"),
                 ("positionZ", "[m]", "float"), ("velocityX", "[m/s]", "float"), ("velocityY", "[m/s]", "float"),
                 ("velocityZ", "[m/s]", "float"), ("tetrahedron", "[1]", "int"))
# Trajectory_3d::status_str (trajectory_3d.cpp:64-83)
TRAJECTORY_STATUS = {0: "NO_INTEGRATION", 1: "NO_FIELD", 2: "STEPPER_LIMIT_PASSED", 3: "ITERATION_LIMIT_PASSED",
                     4: "GEOMETRY_LIMIT_PASSED", 5: "SYSTEM_LIMIT_PASSED"}
_E_CHARGE = np.float32(1.60217733e-19)
_AMU = np.float32(1.660538782e-27)


def trajectory_datasets(events, phases, numbers=None):
    """Datasets of the trajectory stream from tsg_evaporate_traj's events and
    phases: (event index, unique number, time scale, status, charge C, mass
    kg, phases (k, 8), error estimate (6)) per event, the charge and mass in
    float as the event's flight had them (process.cpp:930-935)."""
    out = []
    for e, ph in zip(events, phases):
        charge = np.float32(np.float32(e.ion_charge) * _E_CHARGE)
        mass = np.float32(np.float32(e.ion_mass) * _AMU)
        number = int(numbers[e.node]) if numbers is not None else e.node + 1
        err = np.array(list(e.err_pos) + list(e.err_vel), np.float32)
        out.append((int(e.event_index), number, np.float32(e.time_scale), int(e.traj_status), charge, mass,
                    np.asarray(ph, np.float32), err))
    return out


def _trajectory_header(mode):
    text = "TAPSIM TRAJECTORY DATA\nVERSION " + TRAJECTORY_VERSION + "\n"
    if mode == "BINARY":
        text += "".join("# %s %s (%s)\n" % col for col in _TRAJ_COLUMNS)
    elif mode == "ASCII":
        text += "".join("# %s %s\n" % col[:2] for col in _TRAJ_COLUMNS)
    else:
        raise FormatError("File_Io::LogTrajectory::header(): Invalid output mode!")
    return (text + "\n").encode("latin-1")


def _trajectory_block(d, mode):
    index, number, scale, status, charge, mass, ph, err = d
    scale = np.float32(scale)
    text = "# index = %d\n# unique number = %d\n# time scale = %e\n" % (index, number, float(scale))
    text += "# integrator status = %d => %s\n" % (status, TRAJECTORY_STATUS.get(status, "INVALID"))
    text += "# charge [C] / mass [kg] = %e / %e\n" % (float(charge), float(mass))
    ph = np.asarray(ph, np.float32).reshape(-1, 8)
    t = (ph[:, 0] * scale).astype(np.float32)  # time() * timeScale in float
    if mode == "BINARY":
        rec = np.zeros(len(ph), np.dtype([("t", "<f4"), ("pv", "<f4", 6), ("tet", "<i4")]))
        rec["t"] = t
        rec["pv"] = ph[:, 1:7]
        rec["tet"] = ph[:, 7].astype(np.int32)
        body = ("BINARY %d\n" % len(ph)).encode("latin-1") + rec.tobytes() + b"\n"
    else:
        rows = ["%e\t%s%d\n" % (float(t[k]), "".join("%e\t" % float(v) for v in ph[k, 1:7]), int(ph[k, 7]))
                for k in range(len(ph))]
        body = ("ASCII %d\n" % len(ph) + "".join(rows)).encode("latin-1")
    err = np.asarray(err, np.float32)
    tail = "# final error estimate (position) = (%e/%e/%e)\n" % tuple(float(v) for v in err[:3])
    tail += "# final error estimate (velocity) = (%e/%e/%e)\n\n" % tuple(float(v) for v in err[3:])
    return text.encode("latin-1") + body + tail.encode("latin-1")


def write_trajectories(path, datasets, mode="ASCII", chunk=0):
    """File_Io::LogTrajectory (logTrajectory.cpp:137-299): the header, then
    per event its index, unique number, time scale, integrator status,
    charge / mass, its phases (time scaled by the time scale; tab-separated
    scientific text or packed 32-byte records) and the final error estimate.
    chunk > 0: a new file path.%08u (the event index) whenever the index is a
    multiple of chunk, and at the first event.  Returns the files written."""
    files, cur, name = [], None, None
    try:
        for d in datasets:
            index = int(d[0])
            if cur is None or (chunk > 0 and index % chunk == 0):
                if cur is not None:
                    cur.close()
                name = path if chunk == 0 else "%s.%08u" % (path, index)
                cur = open(name, "wb")
                files.append(name)
                cur.write(_trajectory_header(mode))
            cur.write(_trajectory_block(d, mode))
    finally:
        if cur is not None:
            cur.close()
    return files


# ---------------------------------------------------------------- dump

DUMP_VERSION = "1.0"  # Configuration, Geometry_3d and Grid_3d tables' binary version
GRID_INFINITY_FACTOR = np.float32(1e2)  # Grid_3d::Table::_infinityFactor (grid_3d.h:287)


# a node of the grid block: a 24-byte header (id, number, both potentials,
# charge, three flags, neighbour count) and 8 bytes per neighbour (index,
# coupling): both whole 8-byte units, so the block is an array of them
_GRID_HEAD = np.dtype([("id", "i1"), ("num", "<u4"), ("p0", "<f4"), ("p1", "<f4"), ("q", "<f4"), ("b", "u1"),
                       ("d", "u1"), ("ne", "u1"), ("k", "<i4")], align=False)
_GRID_PAIR = np.dtype([("j", "<i4"), ("c", "<f4")], align=False)


def _grid_records(ids8, numbers, p0, p1, q, fl, ptr, idx, coup):
    n = len(ids8)
    head = np.zeros(n, _GRID_HEAD)
    head["id"], head["num"], head["p0"], head["p1"], head["q"] = ids8, numbers, p0, p1, q
    head["b"], head["d"], head["ne"] = (fl & 1) != 0, (fl & 2) != 0, (fl & 4) != 0
    deg = np.diff(ptr)
    head["k"] = deg
    pair = np.zeros(len(idx), _GRID_PAIR)
    pair["j"], pair["c"] = idx, coup
    units = np.zeros(3 * n + len(idx), "<u8")
    at = 3 * np.arange(n, dtype=np.int64) + ptr[:-1]
    hu = head.view("<u8").reshape(n, 3)
    for w in range(3):
        units[at + w] = hu[:, w]
    row = np.repeat(np.arange(n, dtype=np.int64), deg)
    units[3 * (row + 1) + np.arange(len(idx), dtype=np.int64)] = pair.view("<u8")
    return units


def write_dump(path, last_index, temperature, species, xyz, boundary, assoc, tets, tet_neighbours, extents, ids,
               phi0, phi1, phi_switch, charge, flags, ptr, idx, couplings, numbers=None):
    """File_Io::writeSystem (file_io.cpp:1013-1025): "TAPSIM DUMP INFORMATION",
    "LAST_INDEX <index>", then the binary blocks of the configuration
    (configuration.cpp:162-212), the geometry (geometry_3d.cpp:639-689) and
    the grid (grid_3d.cpp:333-389).  Two fields the reference writes from
    memory it does not own are written as zeros: the two bytes after each
    species' evaporation charge (a short written as a float) and the four
    bytes of the uninitialised relaxation factor (a float written as a double
    whose second half is the potential switch).  species: {id: dict(...)} with
    the fields of read_config; extents: 4 x 3 floats (min, max, constrained
    min, constrained max); flags: bit 1 boundary, 2 Dirichlet, 4 Neumann;
    potentials, charges and couplings are stored as float."""
    out = [b"TAPSIM DUMP INFORMATION\n", ("LAST_INDEX %d\n" % int(last_index)).encode("latin-1")]
    # configuration
    out.append(DUMP_VERSION.encode("latin-1") + b"\0")
    out.append(struct.pack("<fi", float(np.float32(temperature)), len(species)))
    for sid in sorted(species, key=lambda s: np.int8(s)):
        s = species[sid]
        out.append(struct.pack("<b", int(sid)) + str(s.get("name", "")).encode("latin-1") + b"\0")
        out.append(struct.pack("<ff???ffhxxff", s.get("charge_density", 0.0), s.get("epsilon", 1.0),
                               bool(s.get("removable", 0)), bool(s.get("neumann", 0)), bool(s.get("dirichlet", 0)),
                               s.get("phi", 0.0), s.get("mass", 0.0), int(s.get("evap_charge", 0)),
                               s.get("evap_field", 0.0), s.get("evap_energy", 0.0)))
    # geometry
    n = len(xyz)
    out.append(DUMP_VERSION.encode("latin-1") + b"\0")
    out.append(np.asarray(extents, np.float32).reshape(12).tobytes())
    node = np.zeros(n, np.dtype([("p", "<f4", 3), ("b", "u1"), ("t", "<i4")], align=False))
    node["p"] = np.asarray(xyz, np.float32).reshape(n, 3)
    node["b"] = np.asarray(boundary, bool)
    node["t"] = np.asarray(assoc, np.int32)
    out.append(struct.pack("<I", n) + node.tobytes())
    tv = np.asarray(tets, np.int32).reshape(-1, 4)
    tn = np.asarray(tet_neighbours, np.int32).reshape(-1, 4)
    out.append(struct.pack("<I", len(tv)) + np.concatenate([tv, tn], axis=1).astype("<i4").tobytes())
    # grid
    out.append(DUMP_VERSION.encode("latin-1") + b"\0")
    out.append(struct.pack("<f4xii", float(GRID_INFINITY_FACTOR), int(phi_switch), int(phi_switch)))
    out.append(struct.pack("<I", n))
    ptr = np.asarray(ptr, np.int64)
    idx = np.asarray(idx, np.int32)
    coup = np.asarray(couplings, np.float32)
    ids8 = np.asarray(ids).astype(np.int8)
    numbers = np.arange(1, n + 1, dtype=np.uint32) if numbers is None else np.asarray(numbers, np.uint32)
    p0 = np.asarray(phi0, np.float32)
    p1 = np.asarray(phi1, np.float32)
    q = np.asarray(charge, np.float32)
    fl = np.asarray(flags, np.uint8)
    out.append(_grid_records(ids8, numbers, p0, p1, q, fl, ptr, idx, coup).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))


def dump_of_grid(path, grid, mesh, temperature, species, last_index=0, numbers=None):
    """write_dump of a Grid3D's state (its potentials, couplings and charges
    stored as float) with its Mesh"""
    tv, tn = mesh.tets()
    assoc, bnd = mesh.nodeinfo()
    ptr, idx = mesh.csr()
    sw = grid.switches()[0]
    write_dump(path, last_index, temperature, species, mesh.xyz, bnd, assoc, tv, tn, grid.extents(), grid.ids(),
               grid.phi(0), grid.phi(1), sw, grid.charges(), grid.flags(), ptr, idx, grid.couplings(), numbers)


def read_dump(path):
    """File_Io::readSystem (file_io.cpp:1026-1062): the last index and the
    three tables.  As the reference reads them: a species' Dirichlet flag is
    read but not kept (configuration.cpp:135-137: the caller rereads the
    configuration file, tapsim.cpp:161-170), the evaporation charge is the
    short in the first two of its four bytes, the relaxation factor's eight
    bytes end with the potential switch."""
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode("latin-1")
        pos = end + 1
        return s

    def cstr():
        nonlocal pos
        end = data.index(b"\0", pos)
        s = data[pos:end].decode("latin-1")
        pos = end + 1
        return s

    def take(fmt):
        nonlocal pos
        v = struct.unpack_from(fmt, data, pos)
        pos += struct.calcsize(fmt)
        return v

    if not line().startswith("TAPSIM DUMP INFORMATION"):
        raise FormatError("File_Io::readSystem(): wrong file type")
    head = line()
    if not head.startswith("LAST_INDEX"):
        raise FormatError("File_Io::readSystem()")
    last_index = int(head[10:].strip())
    out = {"last_index": last_index}
    if cstr() != DUMP_VERSION:
        raise FormatError("Configuration::Table<<(): wrong version!")
    temperature, count = take("<fi")
    species = {}
    for _ in range(count):
        (sid,) = take("<b")
        name = cstr()
        cd, eps, rem, neu, _dir, phi, mass, q, _pad, ef, ee = take("<ff???ffh2sff")
        species[sid] = dict(name=name, charge_density=cd, epsilon=eps, removable=int(rem), neumann=int(neu),
                            dirichlet=0, phi=phi, mass=mass, evap_charge=q, evap_field=ef, evap_energy=ee)
    out["temperature"], out["species"] = temperature, species
    if cstr() != DUMP_VERSION:
        raise FormatError("Geometry_3d::Table<<(): wrong version!")
    out["extents"] = np.frombuffer(data, np.float32, 12, pos).reshape(4, 3).copy()
    pos += 48
    (n,) = take("<I")
    node = np.frombuffer(data, np.dtype([("p", "<f4", 3), ("b", "u1"), ("t", "<i4")]), n, pos)
    pos += node.nbytes
    out["xyz"], out["boundary"], out["assoc"] = node["p"].copy(), node["b"].astype(bool), node["t"].copy()
    (m,) = take("<I")
    tt = np.frombuffer(data, "<i4", 8 * m, pos).reshape(m, 8)
    pos += tt.nbytes
    out["tets"], out["tet_neighbours"] = tt[:, :4].copy(), tt[:, 4:].copy()
    if cstr() != DUMP_VERSION:
        raise FormatError("Grid_3d::Table<<(): wrong version!")
    inf, _rf, _sw8, sw = take("<f4sii")
    if np.float32(inf) != GRID_INFINITY_FACTOR:
        raise FormatError("Grid_3d::Table::operator<<()")
    out["phi_switch"] = sw
    (nn,) = take("<I")
    # the neighbour counts chain the records: walk them, then take every
    # field at once
    heads = np.zeros(nn, np.int64)
    at = pos
    for i in range(nn):
        heads[i] = at
        at += 24 + 8 * struct.unpack_from("<i", data, at + 20)[0]
    units = np.frombuffer(data, "<u8", (at - pos) // 8, pos)
    u = (heads - pos) // 8
    head = np.stack([units[u], units[u + 1], units[u + 2]], axis=1).copy().view(_GRID_HEAD).reshape(nn)
    deg = head["k"].astype(np.int64)
    ptr = np.zeros(nn + 1, np.int64)
    ptr[1:] = np.cumsum(deg)
    row = np.repeat(np.arange(nn, dtype=np.int64), deg)
    pair = units[3 * (row + 1) + np.arange(ptr[-1], dtype=np.int64)].copy().view(_GRID_PAIR)
    flags = (head["b"] != 0).astype(np.uint8) | ((head["d"] != 0).astype(np.uint8) << 1) | \
        ((head["ne"] != 0).astype(np.uint8) << 2)
    out.update(ids=head["id"].astype(np.int16), numbers=head["num"].copy(), phi0=head["p0"].copy(),
               phi1=head["p1"].copy(), charge=head["q"].copy(), flags=flags, ptr=ptr, idx=pair["j"].copy(),
               couplings=pair["c"].copy())
    return out


# ---------------------------------------------------------------- tapsim.ini

def read_ini(path, delimiter=" = "):
    """File_Io::readInitialization (file_io.cpp:58-94): "KEY = value" lines,
    lines starting with '#' and empty lines skipped, keys and values
    trimmed; a line without the delimiter or a repeated key is an error.
    As in the reference, a last line without its newline is not read."""
    with open(path, "r", encoding="utf-8", errors="surrogateescape") as f:
        text = f.read()
    values = {}
    for line in text.split("\n")[:-1]:
        if not line or line[0] == "#":
            continue
        cut = line.find(delimiter)
        if cut < 0:
            raise FormatError("File_Io::readInitialization(): no %r in %r" % (delimiter, line))
        key, value = line[:cut].strip(), line[cut + len(delimiter):].strip()
        if key in values:
            raise FormatError("File_Io::readInitialization(): duplicate key %s" % key)
        values[key] = value
    return values


# EvaporationOptions::setDefaults (process.cpp:340-440): ini key -> (field of
# tsg_evap_options, parser)
_EVAP_KEYS = (
    ("METHOD_EVAPORATION_PROBABILITY", "prob_mode", int),
    ("METHOD_EVAPORATION_CHOICE", "evap_mode", int),
    ("TRAJECTORY_INTEGRATOR_TYPE", "integrator", int),
    ("TRAJECTORY_STEPPER_TYPE", "stepper", int),
    ("TRAJECTORY_INITIAL_TIME_STEP", "dt0", float),
    ("TRAJECTORY_MINIMUM_TIME_STEP", "dt_min", float),
    ("TRAJECTORY_ERROR_THRESHOLD", "err_limit", float),
    ("VOLTAGE_QUEUE_SIZE", "voltage_queue", int),
    ("LOCAL_RELAXATION_ERROR_THRESHOLD", "relax_threshold", float),
    ("LOCAL_RELAXATION_SHELL_SIZE", "relax_shell", int),
    ("LOCAL_RELAXATION_CYCLE_SIZE", "relax_cycle", int),
    ("LOCAL_RELAXATION_QUEUE_SIZE", "relax_queue", int),
    ("GLOBAL_RELAXATION_STEPS", "relax_global_cycles", int),
    ("REFRESH_RELAXATION_INTERVAL", "refresh_interval", int),
    ("REFRESH_RELAXATION_ERROR_THRESHOLD", "refresh_threshold", float),
    ("REFRESH_RELAXATION_CYCLE_SIZE", "refresh_cycle", int),
    ("REFRESH_RELAXATION_QUEUE_SIZE", "refresh_queue", int),
)


def evap_options_from_ini(values, species):
    """The hot path's evaporation options from tapsim.ini values (read_ini),
    as EvaporationOptions::setDefaults reads them (process.cpp:340-440;
    floats through float, a missing or unreadable key is an error): a dict
    of tsg_evap_options fields, with the vacuum id looked up by
    VACUUM_CELL_IDENTIFIER among the configuration's names, the trajectory
    on, and the thermal start velocity unless TRAJECTORY_NO_INITIAL_VELOCITY."""
    out = {}
    for key, field, kind in _EVAP_KEYS:
        if key not in values:
            raise FormatError("EvaporationOptions::setDefaults(): '%s' missing" % key)
        out[field] = _parse(kind, values[key], key)
    name = values.get("VACUUM_CELL_IDENTIFIER")
    ids = [sid for sid, s in species.items() if s.get("name") == name]
    if name is None or len(ids) != 1:
        raise FormatError("EvaporationOptions::setDefaults(): 'VACUUM_CELL_IDENTIFIER' error!")
    out["vacuum_id"] = ids[0]
    if "TRAJECTORY_NO_INITIAL_VELOCITY" not in values:
        raise FormatError("EvaporationOptions::setDefaults(): 'TRAJECTORY_NO_INITIAL_VELOCITY' missing")
    out["thermal_velocity"] = 0 if _parse(int, values["TRAJECTORY_NO_INITIAL_VELOCITY"], "velocity") else 1
    out["trajectory"] = 1
    return out
