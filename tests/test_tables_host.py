"""Table compiler input checks (CPU): a spaceMap's keywords are checked as spaceMap%init checks them
(Tallies/TallyMaps/Maps1D/spaceMap_class.f90:91-134), with its messages."""
import copy

import pytest

from scone_amd import configs, tables as T


def _slab_with_map(**changes):
    inp, mg = configs.slab_fixed(pop=100, cycles=2, transport="transportOperatorDT")
    inp = copy.deepcopy(inp)
    xax = inp["tally"]["mesh"]["map"]["xax"]
    for k, v in changes.items():
        if v is None:
            xax.pop(k)
        else:
            xax[k] = v
    return inp, mg


def test_lin_and_unstruct_grids_compile():
    inp, mg = _slab_with_map()
    T.compile_problem(inp, mg)
    inp, mg = _slab_with_map(grid="unstruct", bins=[-1.0, 0.0, 0.5, 2.0])
    comp = T.compile_problem(inp, mg)
    assert comp.problem.n_bins == 1 + 3


@pytest.mark.parametrize("changes,message", [
    (dict(grid="log"), "'grid' keyword must be: lin or unstruct"),
    (dict(grid=None), "Keyword 'grid' must be present"),
    (dict(axis=None), "Keyword 'axis' must be present"),
    (dict(axis="w"), "Unrecognised axis: w must be x, y or z"),
])
def test_space_map_keyword_errors(changes, message):
    inp, mg = _slab_with_map(**changes)
    with pytest.raises(ValueError, match=message.replace("(", r"\(").replace(")", r"\)")):
        T.compile_problem(inp, mg)
