"""Every reference citation in the sources (`src/…jl:a-b`, `test/…jl:a`, …) names a file
of the reference tree and a line range inside it.  Runs where the reference tree is
present (the build container); skipped elsewhere."""
import os

import pytest

import citations

REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree absent")
def test_reference_citations_resolve():
    bad = citations.check(REF)
    assert not bad, "\n".join(f"{b[0]}:{b[1]}: {b[2]} ({b[3]})" for b in bad)
