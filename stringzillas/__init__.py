"""The reference's import name: ``import stringzillas as szs`` (python/stringzillas/stringzillas.c:185-327) yields
the MI355X-native package ``stringzillas_amd`` itself, so module state such as ``__capabilities__`` after
``reset_capabilities`` is shared, not copied."""

import sys

import stringzillas_amd as _implementation

sys.modules[__name__] = _implementation
