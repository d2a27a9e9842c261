"""MI355X-native backend for Jacquard's packed Boolean-processor interpreter.

Layers: ``netlist`` (AIG + synthetic generators + AIGPDK Verilog reader),
``script`` (FlattenedScriptV1 compiler/flattener), ``stimulus``, ``vcd``
(reference VCD reader/writer semantics), ``sim`` (plain drop-in runner over
the C ABI), ``batched`` (lane-sliced persistent form), ``events`` (violation
records and timing report), ``costmodel`` (learned partition cost) and
``autotune``.  Native code lives in ``lib/`` (built by ``jacquard_amd.build``).
"""
__all__ = ["netlist", "script", "stimulus", "sim"]
