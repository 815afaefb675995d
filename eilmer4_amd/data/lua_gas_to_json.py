"""Convert an Eilmer gas-model parameter file (a Lua data table) to JSON.

The thermally perfect gas reads, per species, the molecular mass ``M`` and a
CEA/NASA-9 table with ``nsegments``, ``T_break_points``, ``T_blend_ranges`` and
``segment0..segment{n-1}`` (gas/therm_perf_gas.d:63-99 reads ``db.<sp>`` and
its ``thermoCoeffs``; gas/thermo/therm_perf_gas_mix_eos.d:159-173 reads
``<sp>.cea_thermo``; gas/thermo/cea_thermo_curves.d:25-45 reads one table).
This converter understands the subset of Lua those files use: assignments of
numbers, strings and nested table constructors to dotted names.  It executes
nothing from the file.

Usage: python lua_gas_to_json.py <gas-file.lua> <out.json> [table-name]
"""
import json
import re
import sys

_TOKEN = re.compile(r"""
    (?P<ws>\s+|--\[\[.*?\]\]|--[^\n]*)
  | (?P<num>[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)
  | (?P<str>'[^']*'|"[^"]*")
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>[{}=,;.\[\]-])
""", re.VERBOSE | re.DOTALL)


def _tokens(text):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError("cannot tokenise near: %r" % text[pos:pos + 40])
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            continue
        out.append((kind, m.group(kind)))
    return out


class _Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.t[self.i]
        if val is not None and tok[1] != val:
            raise ValueError("expected %r got %r" % (val, tok[1]))
        self.i += 1
        return tok

    def value(self):
        kind, v = self.peek()
        if kind == "num":
            self.take()
            return float(v)
        if kind == "str":
            self.take()
            return v[1:-1]
        if v == "{":
            return self.table()
        if kind == "name" and v in ("true", "false", "nil"):
            self.take()
            return {"true": True, "false": False, "nil": None}[v]
        if v == "-":
            self.take()
            return -self.value()
        raise ValueError("unsupported value token %r" % (v,))

    def table(self):
        self.take("{")
        arr, dct = [], {}
        while self.peek()[1] != "}":
            k0, v0 = self.peek()
            if k0 == "name" and self.peek(1)[1] == "=":
                self.take()
                self.take("=")
                dct[v0] = self.value()
            elif v0 == "[":
                self.take("[")
                key = self.value()
                self.take("]")
                self.take("=")
                dct[key] = self.value()
            else:
                arr.append(self.value())
            if self.peek()[1] in (",", ";"):
                self.take()
        self.take("}")
        if dct and arr:
            dct["__array__"] = arr
            return dct
        return arr if (arr and not dct) else dct

    def chunk(self):
        env = {}
        while self.i < len(self.t):
            kind, v = self.take()
            if kind != "name":
                raise ValueError("expected a name at statement start, got %r" % (v,))
            path = [v]
            while self.peek()[1] == ".":
                self.take(".")
                path.append(self.take()[1])
            self.take("=")
            val = self.value()
            node = env
            for p in path[:-1]:
                node = node.setdefault(p, {})
            node[path[-1]] = val
        return env


def parse_lua_file(path):
    with open(path) as f:
        return _Parser(_tokens(f.read())).chunk()


def thermo_table_to_dict(tab):
    nseg = int(tab["nsegments"])
    return {
        "nsegments": nseg,
        "T_break_points": [float(x) for x in tab["T_break_points"]],
        "T_blend_ranges": [float(x) for x in tab["T_blend_ranges"]],
        "segments": [[float(x) for x in tab["segment%d" % i]] for i in range(nseg)],
    }


def gas_file_to_dict(path, table_name=None):
    env = parse_lua_file(path)
    if table_name:
        return {"model": "CEAThermoCurve", "curve": thermo_table_to_dict(env[table_name])}
    if "IdealGas" in env:
        ig = env["IdealGas"]
        return {"model": "IdealGas", "speciesName": ig["speciesName"],
                "mMass": ig["mMass"], "gamma": ig["gamma"]}
    species = env["species"]
    out = {"model": "ThermallyPerfectGas", "species": species, "M": [], "thermo": []}
    for sp in species:
        rec = env["db"][sp] if "db" in env else env[sp]
        tab = rec.get("thermoCoeffs", rec.get("cea_thermo"))
        out["M"].append(float(rec["M"]))
        out["thermo"].append(thermo_table_to_dict(tab))
    return out


if __name__ == "__main__":
    src, dst = sys.argv[1], sys.argv[2]
    name = sys.argv[3] if len(sys.argv) > 3 else None
    with open(dst, "w") as f:
        json.dump(gas_file_to_dict(src, name), f, indent=1)
        f.write("\n")
