"""Transcribes the CSVR thermostat test data of the reference into
tests/golden/thermostat_goldens.json: the fixture arrays of
test/conftest.py:363-520 (`molecules_with_solvent`: velocities, names) are
read as literals from the file text (nothing is imported or run), and the
mocked random numbers and expected values of test/test_thermostat.py:38-137
are copied by hand below.  Run in a container that has the reference at
/root/reference; the JSON it writes is the committed fixture."""
import ast
import json
import os

REF = "/root/reference/test/conftest.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "thermostat_goldens.json")


def fixture_literals(path, func, names):
    tree = ast.parse(open(path).read())
    fn = next(n for n in ast.walk(tree) if isinstance(n, ast.FunctionDef) and n.name == func)
    out = {}
    for node in ast.walk(fn):
        if isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name in names and isinstance(node.value, ast.Call) and node.value.args:
                out[name] = ast.literal_eval(node.value.args[0])
    return out


lit = fixture_literals(REF, "molecules_with_solvent", {"velocities", "names_"})
data = {
    "_source": ("test/conftest.py:363-520 (molecules_with_solvent: velocities, names) and "
                "test/test_thermostat.py:38-137 (config, mocked random numbers, expected values)"),
    "velocities": lit["velocities"],
    # the fixture appends 15 solvent particles named "D" (conftest.py:548-550);
    # velocities are returned unpermuted, row-aligned with these names
    "names": list(lit["names_"]) + ["D"] * 15,
    "config": {"time_step": 0.032958582578275, "tau": 0.925852989520023, "target_temperature": 310.0,
               "mass": 72.0, "respa_inner": 1, "gas_constant": 0.0083144621},
    "kinetic_energy": 168.45555165866017,
    "kinetic_energy_species": {"A": 55.69986940118755, "B": 59.56037533981311, "C": 7.495327677224677,
                               "D": 45.699979240434836},
    "cases": [
        {"groups": None, "chi_squared": [125.4595634810623], "gaussian": [0.5579657512081987],
         "remove_com": False, "kinetic_energy_new": 171.25339969021243, "work": 2.797848031552252},
        {"groups": [["A"], ["B"], ["C"], ["D"]],
         "chi_squared": [35.43824087713971, 27.57022975113815, 2.024725328228174, 36.50208472031436],
         "gaussian": [-1.752320325907187, 1.099694957420625, 0.6113448515745533, -0.7183266831611322],
         "remove_com": False, "work": -21.670791766960217},
    ],
}
with open(OUT, "w") as fh:
    json.dump(data, fh, indent=1)
print("wrote", OUT, len(data["velocities"]), "particles")
