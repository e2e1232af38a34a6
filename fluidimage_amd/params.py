"""Minimal attribute-tree parameters with the reference's names and defaults.

Stands in for fluiddyn's ParamContainer (``_set_child``, ``_set_attribs``,
attribute access, unknown attributes rejected) for the PIV works:
``WorkPIV.create_default_params()`` returns the same tree as
src/fluidimage/works/piv/multipass.py:47-105 (piv0, piv0.grid, fix,
multipass, mask, series).
"""


class Params:
    def __init__(self, tag="params", attribs=None):
        object.__setattr__(self, "_tag", tag)
        object.__setattr__(self, "_keys", [])
        object.__setattr__(self, "_children", [])
        object.__setattr__(self, "_doc", "")
        if attribs:
            self._set_attribs(attribs)

    def _set_attrib(self, key, value):
        if key not in self._keys:
            self._keys.append(key)
        object.__setattr__(self, key, value)

    def _set_attribs(self, attribs):
        for k, v in attribs.items():
            self._set_attrib(k, v)

    def _set_child(self, tag, attribs=None):
        if tag in self._children:
            child = getattr(self, tag)
            if attribs:
                child._set_attribs(attribs)
            return child
        child = Params(tag, attribs)
        self._children.append(tag)
        object.__setattr__(self, tag, child)
        return child

    def _set_doc(self, doc):
        object.__setattr__(self, "_doc", doc)

    def __setattr__(self, key, value):
        if key in self._keys:
            object.__setattr__(self, key, value)
        elif key in self._children:
            raise AttributeError(f"{key} is a child of {self._tag}, not an attribute")
        else:
            raise AttributeError(
                f"{self._tag} has no parameter {key!r} (known: {self._keys + self._children})"
            )

    def _as_dict(self):
        out = {k: getattr(self, k) for k in self._keys}
        for c in self._children:
            out[c] = getattr(self, c)._as_dict()
        return out

    def __repr__(self):
        return f"<Params {self._tag} {self._as_dict()}>"
