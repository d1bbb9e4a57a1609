"""Tables of results in qsl's on-disk format.

The format is the one Results<T,R>::writeToFile writes
(include/qsl/benchmark/results.tpp:316-355) and the reference's result
files hold (results/np_phase.csv): one "# <line>" per metadata line, an
empty line, the column headings separated by spaces, then one row per line,
integer columns first, every value in the default std::ostream format (six
significant digits, like C's %g).  The methods mirror Results
(include/qsl/benchmark/results.hpp): addMeta, addRow (std::logic_error
cases as ValueError), headings, getMeta, column, mean, variance,
writeToFile; read() parses such a file back.
"""


def _fmt(x):
    """A number as std::ostream << writes it by default."""
    if isinstance(x, int):
        return str(x)
    return "%g" % float(x)


class Results:
    def __init__(self, headings):
        self.head = list(headings)
        self.meta = []
        self.intcols = []
        self.fpcols = []

    def addMeta(self, line):
        self.meta.append(str(line))

    def addRow(self, ints, fps):
        ints, fps = [int(v) for v in ints], [float(v) for v in fps]
        if not self.intcols:
            if len(ints) + len(fps) != len(self.head):
                raise ValueError("Total number of columns does not match number of "
                                 "column headings in first call to addRow")
        elif len(ints) != len(self.intcols[0]):
            raise ValueError("Incorrect number of integer columns in call to addRow")
        elif len(fps) != len(self.fpcols[0]):
            raise ValueError("Incorrect number of floating point columns in call to addRow")
        self.intcols.append(ints)
        self.fpcols.append(fps)

    def headings(self):
        return list(self.head)

    def getMeta(self):
        return list(self.meta)

    def column(self, index):
        if isinstance(index, str):
            index = self.head.index(index)
        if not 0 <= index < len(self.head):
            raise IndexError("Index %d out of range" % index)
        out = []
        for ints, fps in zip(self.intcols, self.fpcols):
            out.append(ints[index] if index < len(ints) else fps[index - len(ints)])
        return out

    def mean(self, index):
        col = self.column(index)
        return sum(col) / len(col)

    def variance(self, index):
        col = self.column(index)
        m = self.mean(index)
        return sum((x - m) ** 2 for x in col) / (len(col) - 1)

    def text(self):
        lines = ["# " + m for m in self.meta]
        lines.append("")
        lines.append(" ".join(self.head))
        for ints, fps in zip(self.intcols, self.fpcols):
            lines.append(" ".join([_fmt(v) for v in ints] + [_fmt(v) for v in fps]))
        return "\n".join(lines) + "\n"

    def writeToFile(self, filename):
        with open(filename, "w") as f:
            f.write(self.text())


def read(path_or_text, nint=1):
    """Parse a results file (or its text).  The first `nint` columns are
    integer columns, the rest floating point."""
    text = path_or_text
    if "\n" not in path_or_text:
        with open(path_or_text) as f:
            text = f.read()
    lines = text.split("\n")
    meta, i = [], 0
    while i < len(lines) and lines[i].startswith("# "):
        meta.append(lines[i][2:])
        i += 1
    if i >= len(lines) or lines[i] != "":
        raise ValueError("results file: expected an empty line after the metadata")
    head = lines[i + 1].split(" ")
    r = Results(head)
    r.meta = meta
    for line in lines[i + 2:]:
        if not line:
            continue
        vals = line.split(" ")
        if len(vals) != len(head):
            raise ValueError("results file: row of %d values under %d headings"
                             % (len(vals), len(head)))
        r.addRow([int(v) for v in vals[:nint]], [float(v) for v in vals[nint:]])
    return r
