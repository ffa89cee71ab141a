"""Name-keyed linear step pipeline (the contract of
rubix/pipeline/abstract_pipeline.py:18-82 and
rubix/pipeline/linear_pipeline.py:56-157): transformers are registered by
their function __name__, the chain is the `depends_on` order of the YAML
`Transformers` block, and the assembled pipeline applies the steps in that
order.  Steps run eagerly on torch's current HIP stream (they are not
jax-traced); `assemble` checks the chain before anything runs.
"""
from __future__ import annotations

from typing import Callable, Dict, List


class LinearTransformerPipeline:
    def __init__(self, cfg: dict, transformers: List[Callable]):
        self.config = cfg
        self.transformers: Dict[str, Callable] = {}
        self._names: List[str] = []
        self._pipeline: List[Callable] = []
        for t in transformers:
            self.register_transformer(t)
        self.assemble()

    def register_transformer(self, fn: Callable):
        if fn.__name__ in self.transformers:
            raise ValueError("A transformer of this name is already present")
        self.transformers[fn.__name__] = fn

    @property
    def pipeline(self) -> dict:
        return dict(zip(self._names, self._pipeline))

    def build_pipeline(self):
        """Order the configured transformers along their depends_on links:
        exactly one start node, each node followed by at most one other."""
        if not self.transformers:
            raise RuntimeError("Cannot build a pipeline without registered transformers")
        nodes = self.config["Transformers"]
        starts = [k for k, v in nodes.items() if v.get("depends_on") is None]
        if len(starts) != 1:
            raise ValueError("A linear pipeline needs exactly one transformer without depends_on")
        children = {}
        for k, v in nodes.items():
            d = v.get("depends_on")
            if d is None:
                continue
            if d in children:
                raise ValueError(f"Transformer {d} has more than one successor; not linear")
            if d not in nodes:
                raise ValueError(f"Transformer {k} depends on unknown {d}")
            children[d] = k
        order, node = [], starts[0]
        while node is not None:
            order.append(node)
            node = children.get(node)
        if len(order) != len(nodes):
            raise ValueError("The transformer chain is not connected")
        self._names, self._pipeline = [], []
        for key in order:
            fname = nodes[key]["name"]
            if fname not in self.transformers:
                raise ValueError(f"Transformer {fname} is not registered")
            self._names.append(key)
            self._pipeline.append(self.transformers[fname])

    def assemble(self):
        self.build_pipeline()

    def apply(self, data, *args, **kwargs):
        for fn in self._pipeline:
            data = fn(data, *args, **kwargs)
        return data

    __call__ = apply
