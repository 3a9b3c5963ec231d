"""String labels with primes (mirrors tncontract/label.py:17-92).

A primed label remembers the label it was primed from, so primes can be
removed again; ``unique_label`` returns a uuid4 string.
"""
from __future__ import annotations

import uuid

__all__ = ["prime_label", "unprime_label", "noprime_label", "prime_level", "unique_label"]


class Label(str):
    """A ``str`` that keeps a link to the label it was primed from."""

    def __new__(cls, value, parent=None):
        return str.__new__(cls, value)

    def __init__(self, value, parent=None):
        if parent is None and isinstance(value, Label):
            parent = value._parent
        self._parent = parent

    @property
    def parent(self):
        return self._parent

    @property
    def origin(self):
        node = self
        while getattr(node, "parent", None) is not None:
            node = node.parent
        return node

    @property
    def parents(self):
        level, node = 0, self
        while getattr(node, "parent", None) is not None:
            node = node.parent
            level += 1
        return level


def prime_label(label, prime="'"):
    return Label(str(label) + prime, parent=label)


def unprime_label(label, prime="'"):
    parent = getattr(label, "parent", None)
    if not isinstance(label, Label):
        raise ValueError("label is not primed")
    if parent is not None and str(parent) + prime == label:
        return parent
    raise ValueError("label is not primed with \"" + prime + "\"")


def noprime_label(label):
    return label.origin if isinstance(label, Label) else label


def prime_level(label):
    return label.parents if isinstance(label, Label) else 0


def unique_label():
    return str(uuid.uuid4())
