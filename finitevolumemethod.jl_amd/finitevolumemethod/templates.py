"""Linear templates (reference src/specific_problems/*.jl)."""


class AbstractFVMTemplate:
    """abstract_templates.jl:1-38."""

    def solve(self, alg=None, **kwargs):
        raise NotImplementedError
