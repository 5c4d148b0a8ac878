"""OpenQASM 3.0 parser over the language grammar, with a constant-expression
evaluator, and its emitter
(src/org/soulspace/qclojure/application/format/qasm3_parser.clj,
resources/openqasm/qasm3.ebnf).

The reference parses with Instaparse over ``qasm3.ebnf`` (whitespace allowed
between any two terminals) and then keeps only two kinds of top-level
statement: the first ``qubit[n] name;`` (its size) and gate calls without
modifiers whose name the operation registry knows (aliases resolved).
Everything else that the grammar accepts — classical declarations,
measurements, resets, barriers, gate definitions, control flow — is checked
for syntax and dropped; text the grammar rejects raises ``QasmParseError``.
This module restates that with a hand-written recursive-descent parser over
the same productions (ordered choice where the grammar is ambiguous) and the
reference's evaluation rules:

* gate parameters are constant expressions with OpenQASM 3 precedence,
  integer arithmetic on 64-bit longs (overflow is an error), ``/`` always in
  doubles, unary ``-``/``+`` and ``**`` always in doubles, constants ``pi``,
  ``tau``, ``euler`` and the built-in functions sin … rotr;
* a qubit index or register size is the first integer literal inside the
  brackets (``q[1+2]`` is qubit 1, ``q[i]`` is no operand);
* an identifier may not begin with a reserved word (``index`` starts with
  ``in``), as in the grammar's negative lookahead.

The circuit is named "Parsed QASM3 Circuit"; ``result-specs`` come from the
``#pragma qclojure result`` lines.
"""
from __future__ import annotations

import logging
import math
import re

from . import circuit as qc
from .jvmtext import clj_str, java_div, java_parse_double, java_parse_int, JvmNumberFormatError
from .qasm3 import QasmError, java_split, java_trim, split_lines, PRAGMA_TARGET, clj_name

log = logging.getLogger(__name__)


class QasmParseError(QasmError):
    """Text the grammar does not accept (the reference's ``Parse error``)."""


# ---- result pragmas (qasm3_parser.clj:19-64) --------------------------------------
_DIGITS_FULL = re.compile(r"\d+", re.ASCII)


def parse_result_pragma(line: str):
    """parse-result-pragma (qasm3_parser.clj:19-43): as the line converter's,
    except that list elements made of digits read as integers."""
    if not line.startswith("#pragma qclojure result"):
        return None
    parts = java_split(line, r"\s+")
    if len(parts) <= 3:
        return None
    params: dict = {}
    for pair in parts[4:]:
        kv = java_split(pair, "=", 2)
        if len(kv) < 2:
            raise QasmError(f"NullPointerException: pragma parameter '{pair}' has no value")
        k, v = kv
        if _DIGITS_FULL.fullmatch(v):
            params[k] = java_parse_int(v, 64)
        elif "," in v:
            params[k] = [java_parse_int(x, 64) if _DIGITS_FULL.fullmatch(x) else java_trim(x)
                         for x in java_split(v, ",")]
        else:
            params[k] = v
    return [parts[3], params]


def collect_result_specs_from_qasm(lines) -> dict:
    """collect-result-specs-from-qasm (qasm3_parser.clj:45-64)."""
    specs: dict = {}
    for line in lines:
        if not java_trim(line).startswith("#pragma qclojure result"):
            continue
        parsed = parse_result_pragma(line)
        if parsed is None:
            continue
        rtype, params = parsed
        existing = specs.get(rtype)
        if existing is not None and rtype in ("expectation", "variance", "sample"):
            e = dict(existing)
            e["observables"] = list(e.get("observables") or []) + [params.get("observable")]
            e["targets"] = list(e.get("targets") or []) + [params.get("target")]
            specs[rtype] = e
        else:
            specs[rtype] = params
    return specs


# ---- comments (qasm3_parser.clj:156-175) ------------------------------------------------
def remove_line_comments(code: str) -> str:
    """remove-line-comments: ``//`` to the end of each line."""
    return re.sub(r"//.*?(?=\r?\n|$)", "", code)


def remove_block_comments(code: str) -> str:
    """remove-block-comments: ``/* ... */`` (across lines)."""
    return re.sub(r"(?s)/\*.*?\*/", "", code)


def preprocess_qasm(code: str) -> str:
    """preprocess-qasm: block comments, then line comments."""
    return remove_line_comments(remove_block_comments(code))


# ---- grammar (resources/openqasm/qasm3.ebnf) ---------------------------------------------------
RESERVED = ("qubit", "bit", "int", "uint", "float", "angle", "bool", "duration", "stretch", "complex",
            "include", "gate", "def", "defcal", "extern", "box", "let", "break", "continue",
            "if", "else", "end", "return", "for", "while", "in", "switch", "case", "default",
            "input", "output", "const", "readonly", "mutable", "qreg", "creg", "void",
            "gphase", "inv", "pow", "ctrl", "negctrl", "delay", "reset", "measure", "barrier",
            "nop", "true", "false")
CLASSICAL_TYPES = ("bit", "int", "uint", "float", "angle", "bool", "duration", "stretch", "complex")
_WS = re.compile(r"[ \t\n\x0b\f\r]+")
_IDENT = re.compile(r"[A-Za-z_][A-Za-z0-9_]*")
# Longest match wins; on a tie the earlier form (digits alone read as an integer).
_NUMBER_FORMS = (
    ("int", re.compile(r"[0-9]+")),
    ("float", re.compile(r"[0-9]+(\.[0-9]*)?([eE][+-]?[0-9]+)?")),
    ("timing", re.compile(r"[0-9]+(\.[0-9]*)?(dt|ns|us|µs|ms|s)")),
    ("imag", re.compile(r"[0-9]+(\.[0-9]*)?[eE]?[+-]?[0-9]*im")),
)
_STRING = re.compile(r"\"[^\"]*\"|'[^']*'")
_HWQ = re.compile(r"\$[0-9]+")
_COMPOUND = ("<<=", ">>=", "**=", "+=", "-=", "*=", "/=", "&=", "|=", "^=", "%=")


class _Fail(Exception):
    pass


class _Parser:
    """Recursive descent over the grammar's productions. Expression nodes are
    tuples: ("int", text), ("float", text), ("bool", v), ("ident", name),
    ("hwq", text), ("lit", kind, text), ("call", name, args), ("unary", op, x),
    ("pow", a, b), ("bin", op, a, b), ("paren", x), ("range", a, b, c),
    ("set", items), ("array", items)."""

    def __init__(self, text: str):
        self.s = text
        self.i = 0
        self.decls: list = []   # ("qdecl", designator expr or None)
        self.calls: list = []   # ("call", name or None, params, operands)

    # -- lexical helpers
    def ws(self):
        m = _WS.match(self.s, self.i)
        if m:
            self.i = m.end()

    def at_end(self) -> bool:
        self.ws()
        return self.i >= len(self.s)

    def peek_lit(self, lit: str) -> bool:
        self.ws()
        return self.s.startswith(lit, self.i)

    def lit(self, lit: str) -> None:
        if not self.peek_lit(lit):
            raise _Fail(lit)
        self.i += len(lit)

    def opt_lit(self, lit: str) -> bool:
        if self.peek_lit(lit):
            self.i += len(lit)
            return True
        return False

    def op(self, choices) -> str | None:
        """Longest operator among ``choices`` at the cursor."""
        self.ws()
        best = None
        for c in choices:
            if self.s.startswith(c, self.i) and (best is None or len(c) > len(best)):
                best = c
        if best is not None:
            self.i += len(best)
        return best

    def identifier(self) -> str:
        self.ws()
        for w in RESERVED:
            if self.s.startswith(w, self.i):
                raise _Fail("reserved")
        m = _IDENT.match(self.s, self.i)
        if not m:
            raise _Fail("identifier")
        self.i = m.end()
        return m.group(0)

    def try_(self, fn, *args):
        save = self.i, len(self.decls), len(self.calls)
        try:
            return True, fn(*args)
        except _Fail:
            self.i = save[0]
            del self.decls[save[1]:]
            del self.calls[save[2]:]
            return False, None

    def rest_of_line(self) -> str:
        j = self.s.find("\n", self.i)
        j = len(self.s) if j < 0 else j
        t, self.i = self.s[self.i:j], j
        return t

    # -- program
    def program(self):
        save = self.i
        if self.opt_lit("OPENQASM"):
            self.ws()
            m = re.compile(r"[0-9]+(\.[0-9]+)?").match(self.s, self.i)
            if not m:
                raise _Fail("version")
            self.i = m.end()
            self.lit(";")
        else:
            self.i = save
        while not self.at_end():
            self.statement(top=True)

    def statement(self, top: bool = False):
        self.ws()
        start, dstart = len(self.calls), len(self.decls)
        alternatives = (self.pragma, self.annotation, self.include_statement, self.gate_statement,
                        self.quantum_declaration, self.classical_declaration, self.alias_declaration,
                        self.const_declaration, self.reset_statement, self.measure_statement,
                        self.barrier_statement, self.delay_statement, self.box_statement,
                        self.extern_statement, self.def_statement, self.defcal_statement,
                        self.for_statement, self.if_statement, self.while_statement,
                        self.return_statement, self.switch_statement, self.nop_statement,
                        self.gate_call_statement, self.assignment_statement, self.expression_statement)
        for alt in alternatives:
            ok, _ = self.try_(alt)
            if ok:
                if not top:   # statements inside a scope are not the program's
                    del self.calls[start:]
                    del self.decls[dstart:]
                return
        raise _Fail("statement")

    def pragma(self):
        self.ws()
        if self.s.startswith("#pragma", self.i):
            self.i += 7
        elif self.s.startswith("pragma", self.i):
            self.i += 6
        else:
            raise _Fail("pragma")
        self.ws()
        self.rest_of_line()

    def annotation(self):
        self.lit("@")
        self.identifier()
        while True:
            ok, _ = self.try_(self._dot_ident)
            if not ok:
                break
        self.rest_of_line()

    def _dot_ident(self):
        self.lit(".")
        self.identifier()

    def string_literal(self):
        self.ws()
        m = _STRING.match(self.s, self.i)
        if not m:
            raise _Fail("string")
        self.i = m.end()
        return m.group(0)

    def include_statement(self):
        self.lit("include")
        self.string_literal()
        self.lit(";")

    def identifier_list(self):
        out = [self.identifier()]
        while True:
            save = self.i
            if not self.opt_lit(","):
                break
            ok, v = self.try_(self.identifier)
            if not ok:
                self.i = save
                self.opt_lit(",")
                break
            out.append(v)
        return out

    def gate_statement(self):
        self.lit("gate")
        self.identifier()
        if self.opt_lit("("):
            if not self.peek_lit(")"):
                self.identifier_list()
            self.lit(")")
        self.identifier_list()
        self.scope()

    def designator(self):
        self.lit("[")
        e = self.expression()
        self.lit("]")
        return e

    def quantum_declaration(self):
        self.lit("qubit")
        ok, d = self.try_(self.designator)
        self.identifier()
        self.lit(";")
        self.decls.append(("qdecl", d if ok else None))

    def classical_type(self):
        self.ws()
        for t in CLASSICAL_TYPES:
            if self.s.startswith(t, self.i):
                self.i += len(t)
                return t
        raise _Fail("classical type")

    def classical_declaration(self):
        self.classical_type()
        self.try_(self.designator)
        self.identifier()
        if self.opt_lit("="):
            self.declaration_expression()
        self.lit(";")

    def declaration_expression(self):
        for alt in (self.array_literal, self.expression, self.measure_expression):
            ok, _ = self.try_(alt)
            if ok:
                return
        raise _Fail("declaration expression")

    def scalar_type(self):
        self.ws()
        for t in ("bit", "int", "uint", "float", "angle"):
            if self.s.startswith(t, self.i):
                self.i += len(t)
                self.try_(self.designator)
                return
        for t in ("bool", "duration", "stretch"):
            if self.s.startswith(t, self.i):
                self.i += len(t)
                return
        if self.s.startswith("complex", self.i):
            self.i += 7
            save = self.i
            if self.opt_lit("["):
                ok, _ = self.try_(self.scalar_type)
                if ok and self.opt_lit("]"):
                    return
                self.i = save
            return
        raise _Fail("scalar type")

    def alias_declaration(self):
        self.lit("let")
        self.identifier()
        self.lit("=")
        self.expression()
        self.lit(";")

    def const_declaration(self):
        self.lit("const")
        self.scalar_type()
        self.identifier()
        self.lit("=")
        self.declaration_expression()
        self.lit(";")

    def index_operator(self):
        self.lit("[")
        ok, v = self.try_(self.set_expression)
        if ok:
            self.lit("]")
            return v
        first = self._expr_or_range()
        while True:
            save = self.i
            if not self.opt_lit(","):
                break
            ok, _ = self.try_(self._expr_or_range)
            if not ok:
                self.i = save
                self.opt_lit(",")
                break
        self.lit("]")
        return first

    def _expr_or_range(self):
        ok, v = self.try_(self._range_then_check)
        if ok:
            return v
        return self.expression()

    def _range_then_check(self):
        v = self.range_expression()
        if not (self.peek_lit("]") or self.peek_lit(",")):
            raise _Fail("range")
        return v

    def indexed_identifier(self):
        name = self.identifier()
        idx = []
        while True:
            ok, v = self.try_(self.index_operator)
            if not ok:
                break
            idx.append(v)
        return ("indexed", name, idx)

    def gate_operand(self):
        self.ws()
        m = _HWQ.match(self.s, self.i)
        if m:
            self.i = m.end()
            return ("hwq", m.group(0))
        return self.indexed_identifier()

    def gate_operand_list(self):
        out = [self.gate_operand()]
        while True:
            save = self.i
            if not self.opt_lit(","):
                break
            ok, v = self.try_(self.gate_operand)
            if not ok:
                self.i = save
                self.opt_lit(",")
                break
            out.append(v)
        return out

    def reset_statement(self):
        self.lit("reset")
        self.gate_operand()
        self.lit(";")

    def measure_expression(self):
        self.lit("measure")
        self.gate_operand()

    def measure_statement(self):
        self.measure_expression()
        if self.opt_lit("->"):
            self.indexed_identifier()
        self.lit(";")

    def barrier_statement(self):
        self.lit("barrier")
        self.try_(self.gate_operand_list)
        self.lit(";")

    def delay_statement(self):
        self.lit("delay")
        self.designator()
        self.try_(self.gate_operand_list)
        self.lit(";")

    def nop_statement(self):
        self.lit("nop")
        self.try_(self.gate_operand_list)
        self.lit(";")

    def box_statement(self):
        self.lit("box")
        self.try_(self.designator)
        self.scope()

    def scope(self):
        self.lit("{")
        while not self.peek_lit("}"):
            if self.at_end():
                raise _Fail("scope")
            self.statement_or_scope()
        self.lit("}")

    def statement_or_scope(self):
        ok, _ = self.try_(self.statement)
        if not ok:
            self.scope()

    def return_signature(self):
        self.lit("->")
        self.scalar_type()

    def array_reference_type(self):
        self.ws()
        if not (self.opt_lit("readonly") or self.opt_lit("mutable")):
            raise _Fail("array ref")
        self.lit("array")
        self.lit("[")
        self.scalar_type()
        self.lit(",")
        if self.opt_lit("#dim"):
            self.lit("=")
            self.expression()
        else:
            self.expression_list()
        self.lit("]")

    def extern_argument(self):
        for alt in (self.scalar_type, self.array_reference_type, self._creg_arg):
            ok, _ = self.try_(alt)
            if ok:
                return
        raise _Fail("extern argument")

    def _creg_arg(self):
        self.lit("creg")
        self.try_(self.designator)

    def _list_of(self, item):
        item()
        while True:
            save = self.i
            if not self.opt_lit(","):
                break
            ok, _ = self.try_(item)
            if not ok:
                self.i = save
                self.opt_lit(",")
                break

    def extern_statement(self):
        self.lit("extern")
        self.identifier()
        self.lit("(")
        if not self.peek_lit(")"):
            self._list_of(self.extern_argument)
        self.lit(")")
        self.try_(self.return_signature)
        self.lit(";")

    def argument_definition(self):
        def scalar_arg():
            self.scalar_type()
            self.identifier()

        def qubit_arg():
            self.lit("qubit")
            self.try_(self.designator)
            self.identifier()

        def reg_arg():
            if not (self.opt_lit("creg") or self.opt_lit("qreg")):
                raise _Fail("reg")
            self.identifier()
            self.try_(self.designator)

        def array_arg():
            self.array_reference_type()
            self.identifier()

        for alt in (scalar_arg, qubit_arg, reg_arg, array_arg):
            ok, _ = self.try_(alt)
            if ok:
                return
        raise _Fail("argument definition")

    def def_statement(self):
        self.lit("def")
        self.identifier()
        self.lit("(")
        if not self.peek_lit(")"):
            self._list_of(self.argument_definition)
        self.lit(")")
        self.try_(self.return_signature)
        self.scope()

    def defcal_statement(self):
        self.lit("defcal")
        self.ws()
        if not (self.opt_lit("measure") or self.opt_lit("reset") or self.opt_lit("delay")):
            self.identifier()
        if self.opt_lit("("):
            if not self.peek_lit(")"):
                def arg():
                    ok, _ = self.try_(self.expression)
                    if not ok:
                        self.argument_definition()
                self._list_of(arg)
            self.lit(")")

        def operand():
            self.ws()
            m = _HWQ.match(self.s, self.i)
            if m:
                self.i = m.end()
            else:
                self.identifier()
        self._list_of(operand)
        self.try_(self.return_signature)
        self.lit("{")
        j = self.s.find("}", self.i)
        if j < 0:
            raise _Fail("defcal body")
        self.i = j + 1

    def for_statement(self):
        self.lit("for")
        self.scalar_type()
        self.identifier()
        self.lit("in")
        ok, _ = self.try_(self.set_expression)
        if not ok:
            save = self.i
            if self.opt_lit("["):
                ok, _ = self.try_(self.range_expression)
                if ok and self.opt_lit("]"):
                    pass
                else:
                    self.i = save
                    self.expression()
            else:
                self.expression()
        self.statement_or_scope()

    def if_statement(self):
        self.lit("if")
        self.lit("(")
        self.expression()
        self.lit(")")
        self.statement_or_scope()
        save = self.i
        if self.opt_lit("else"):
            ok, _ = self.try_(self.statement_or_scope)
            if not ok:
                self.i = save

    def while_statement(self):
        self.lit("while")
        self.lit("(")
        self.expression()
        self.lit(")")
        self.statement_or_scope()

    def return_statement(self):
        self.lit("return")
        if not self.peek_lit(";"):
            ok, _ = self.try_(self.expression)
            if not ok:
                self.measure_expression()
        self.lit(";")

    def switch_statement(self):
        self.lit("switch")
        self.lit("(")
        self.expression()
        self.lit(")")
        self.lit("{")
        while not self.peek_lit("}"):
            if self.opt_lit("case"):
                self.expression_list()
                self.scope()
            elif self.opt_lit("default"):
                self.scope()
            else:
                raise _Fail("switch item")
        self.lit("}")

    def gate_modifier(self):
        self.ws()
        if self.opt_lit("inv"):
            self.lit("@")
            return
        if self.opt_lit("pow"):
            self.lit("(")
            self.expression()
            self.lit(")")
            self.lit("@")
            return
        if self.opt_lit("negctrl") or self.opt_lit("ctrl"):
            if self.opt_lit("("):
                self.expression()
                self.lit(")")
            self.lit("@")
            return
        raise _Fail("modifier")

    def gate_call_statement(self):
        modifiers = 0
        while True:
            ok, _ = self.try_(self.gate_modifier)
            if not ok:
                break
            modifiers += 1
        if self.opt_lit("gphase"):
            name = None
        else:
            name = self.identifier()
        params = None
        if self.opt_lit("("):
            params = [] if self.peek_lit(")") else self.expression_list()
            self.lit(")")
        self.try_(self.designator)
        ok, operands = self.try_(self.gate_operand_list)
        self.lit(";")
        self.calls.append(("call", None if modifiers else name, params or [], operands if ok else []))

    def assignment_statement(self):
        self.indexed_identifier()
        if self.op(("=",) + _COMPOUND) is None or self.s[self.i - 1:self.i + 1] == "==":
            raise _Fail("assignment")
        ok, _ = self.try_(self.expression)
        if not ok:
            self.measure_expression()
        self.lit(";")

    def expression_statement(self):
        self.expression()
        self.lit(";")

    # -- expressions
    def expression_list(self):
        out = [self.expression()]
        while True:
            save = self.i
            if not self.opt_lit(","):
                break
            ok, v = self.try_(self.expression)
            if not ok:
                self.i = save
                self.opt_lit(",")
                break
            out.append(v)
        return out

    def expression(self):
        return self.logical_or()

    def _left(self, sub, ops):
        x = sub()
        while True:
            save = self.i
            o = self.op(ops)
            if o is None:
                return x
            ok, y = self.try_(sub)
            if not ok:
                self.i = save
                return x
            x = ("bin", o, x, y)

    def logical_or(self):
        return self._left(self.logical_and, ("||",))

    def logical_and(self):
        return self._left(self.bitwise_or, ("&&",))

    def bitwise_or(self):
        return self._left_excl(self.bitwise_xor, "|", "||")

    def bitwise_xor(self):
        return self._left(self.bitwise_and, ("^",))

    def bitwise_and(self):
        return self._left_excl(self.equality, "&", "&&")

    def _left_excl(self, sub, single, double):
        x = sub()
        while True:
            save = self.i
            self.ws()
            if self.s.startswith(single, self.i) and not self.s.startswith(double, self.i):
                self.i += 1
                ok, y = self.try_(sub)
                if ok:
                    x = ("bin", single, x, y)
                    continue
            self.i = save
            return x

    def equality(self):
        return self._left(self.comparison, ("==", "!="))

    def comparison(self):
        x = self.additive()
        while True:
            save = self.i
            o = self.op((">", "<", ">=", "<="))
            if o is None:
                return x
            ok, y = self.try_(self.additive)
            if not ok:
                self.i = save
                return x
            x = ("bin", o, x, y)

    def additive(self):
        return self._left(self.multiplicative, ("+", "-"))

    def multiplicative(self):
        x = self.power()
        while True:
            save = self.i
            self.ws()
            if self.s.startswith("**", self.i):
                return x
            o = self.op(("*", "/", "%"))
            if o is None:
                return x
            ok, y = self.try_(self.power)
            if not ok:
                self.i = save
                return x
            x = ("bin", o, x, y)

    def power(self):
        a = self.unary()
        save = self.i
        if self.opt_lit("**"):
            ok, b = self.try_(self.power)
            if ok:
                return ("pow", a, b)
            self.i = save
        return a

    def unary(self):
        self.ws()
        c = self.s[self.i:self.i + 1]
        if c in ("~", "!", "-", "+") and not (c == "!" and self.s.startswith("!=", self.i)):
            save = self.i
            self.i += 1
            ok, x = self.try_(self.unary)
            if ok:
                return ("unary", c, x)
            self.i = save
        return self.primary()

    def primary(self):
        self.ws()
        if self.opt_lit("("):
            e = self.expression()
            self.lit(")")
            return ("paren", e)
        ok, v = self.try_(self.call_expression)
        if ok:
            return v
        ok, v = self.try_(self.literal)
        if ok:
            return v
        ok, v = self.try_(self.identifier)
        if ok:
            return ("ident", v)
        m = _HWQ.match(self.s, self.i)
        if m:
            self.i = m.end()
            return ("hwq", m.group(0))
        raise _Fail("primary")

    def call_expression(self):
        name = self.identifier()
        self.lit("(")
        args = [] if self.peek_lit(")") else self.expression_list()
        self.lit(")")
        return ("call", name, args)

    def literal(self):
        self.ws()
        best = None
        for kind, rx in _NUMBER_FORMS:
            m = rx.match(self.s, self.i)
            if m and (best is None or m.end() > best[1].end()):
                best = (kind, m)
        if best is not None:
            kind, m = best
            self.i = m.end()
            if kind in ("int", "float"):
                return (kind, m.group(0))
            return ("lit", kind, m.group(0))
        for b in ("true", "false"):
            if self.s.startswith(b, self.i):
                self.i += len(b)
                return ("bool", b == "true")
        ok, v = self.try_(self.string_literal)
        if ok:
            return ("lit", "bitstring", v)
        raise _Fail("literal")

    def range_expression(self):
        a = b = c = None
        ok, a = self.try_(self.expression)
        self.lit(":")
        ok, b = self.try_(self.expression)
        save = self.i
        if self.opt_lit(":"):
            ok, c = self.try_(self.expression)
            if not ok:
                self.i = save
        return ("range", a, b, c)

    def set_expression(self):
        self.lit("{")
        items = self.expression_list()
        self.lit("}")
        return ("set", items)

    def array_literal(self):
        self.lit("{")
        items = []
        if not self.peek_lit("}"):
            def item():
                ok, v = self.try_(self.array_literal)
                items.append(v if ok else self.expression())
            self._list_of(item)
        self.lit("}")
        return ("array", items)


# ---- constant-expression evaluation (qasm3_parser.clj:198-512) --------------------------------
CONSTANTS = {"pi": math.pi, "π": math.pi, "tau": 2.0 * math.pi, "τ": 2.0 * math.pi,
             "euler": math.e, "ℇ": math.e}
_LMIN, _LMAX = -(1 << 63), (1 << 63) - 1


class _EvalError(Exception):
    pass


def _long(v: int) -> int:
    if not _LMIN <= v <= _LMAX:
        raise _EvalError("integer overflow")
    return v


def to_bool(v) -> bool:
    """to-bool (qasm3_parser.clj:207-213)."""
    if isinstance(v, bool):
        return v
    if isinstance(v, (int, float)):
        return not (float(v) == 0.0)
    raise _EvalError("Cannot cast to bool")


def _jfloor(x: float) -> float:
    return float(math.floor(x)) if math.isfinite(x) else x


def _jceil(x: float) -> float:
    return float(math.ceil(x)) if math.isfinite(x) else x


def to_long(v) -> int:
    """to-long (qasm3_parser.clj:215-223): floats are floored, then cast with
    the JVM's range check (NaN casts to 0)."""
    if v is None:
        return 0
    if isinstance(v, bool):
        return 1 if v else 0
    if isinstance(v, int):
        return v
    if isinstance(v, float):
        if math.isnan(v):
            return 0
        f = _jfloor(v)
        if f < -9.223372036854776e18 or f > 9.223372036854776e18 or math.isinf(f):
            raise _EvalError("Value out of range for long")
        return max(_LMIN, min(_LMAX, int(f)))
    raise _EvalError("Cannot cast to integer")


def to_double(v) -> float:
    """to-double (qasm3_parser.clj:225-232)."""
    if v is None:
        return 0.0
    if isinstance(v, bool):
        return 1.0 if v else 0.0
    if isinstance(v, (int, float)):
        return float(v)
    raise _EvalError("Cannot cast to double")


def _is_double(v) -> bool:
    return isinstance(v, float)


def _signed64(x: int) -> int:
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >> 63 else x


def java_pow(a: float, b: float) -> float:
    """``Math.pow`` special cases (IEEE pow)."""
    if b == 0.0:
        return 1.0
    if math.isnan(a) or math.isnan(b):
        return math.nan
    try:
        return math.pow(a, b)
    except ValueError:
        if a == 0.0 and b < 0:
            odd = float(b).is_integer() and int(b) % 2 == 1
            return -math.inf if (odd and math.copysign(1.0, a) < 0) else math.inf
        return math.nan
    except OverflowError:
        odd = float(b).is_integer() and int(b) % 2 == 1
        return -math.inf if (a < 0 and odd) else math.inf


def _jmath(fn, x: float) -> float:
    try:
        return fn(x)
    except ValueError:
        return math.nan
    except OverflowError:
        return math.inf


def _jlog(x: float) -> float:
    if math.isnan(x) or x < 0:
        return math.nan
    if x == 0.0:
        return -math.inf
    return math.log(x)


def _jsqrt(x: float) -> float:
    if math.isnan(x) or x < 0:
        return math.nan
    return math.sqrt(x)


def _double_mod(a, b) -> float:
    ad, bd = to_double(a), to_double(b)
    q = java_div(ad, bd)
    f = _jfloor(q)
    prod = f * bd if not (math.isinf(f) and bd == 0.0) else math.nan
    return ad - prod


def _clj_mod(a: int, b: int) -> int:
    """``(mod a b)`` on longs: floored modulus (the sign of the divisor)."""
    if b == 0:
        raise _EvalError("Divide by zero")
    return a % b


def apply_function(fname: str, args: list):
    """apply-function (qasm3_parser.clj:297-339)."""
    f = fname.lower()
    a0 = args[0] if args else None
    a1 = args[1] if len(args) > 1 else None
    if f == "sin":
        return _jmath(math.sin, to_double(a0))
    if f == "cos":
        return _jmath(math.cos, to_double(a0))
    if f == "tan":
        return _jmath(math.tan, to_double(a0))
    if f == "arcsin":
        return _jmath(math.asin, to_double(a0))
    if f == "arccos":
        return _jmath(math.acos, to_double(a0))
    if f == "arctan":
        return _jmath(math.atan, to_double(a0))
    if f == "exp":
        return _jmath(math.exp, to_double(a0))
    if f == "sqrt":
        return _jsqrt(to_double(a0))
    if f == "floor":
        return _jfloor(to_double(a0))
    if f in ("ceiling", "ceil"):
        return _jceil(to_double(a0))
    if f == "log":
        if len(args) == 1:
            return _jlog(to_double(a0))
        if len(args) == 2:
            return java_div(_jlog(to_double(a0)), _jlog(to_double(a1)))
        raise _EvalError("log expects 1 or 2 args")
    if f == "pow":
        return java_pow(to_double(a0), to_double(a1))
    if f == "mod":
        if isinstance(a0, int) and not isinstance(a0, bool) and isinstance(a1, int) and not isinstance(a1, bool):
            return _clj_mod(to_long(a0), to_long(a1))
        return _double_mod(a0, a1)
    if f == "popcount":
        return bin(to_long(a0) & ((1 << 64) - 1)).count("1")
    if f in ("rotl", "rotr"):
        v, n = to_long(a0), to_long(a1) % 64
        u = v & ((1 << 64) - 1)
        if f == "rotl":
            r = (u << n) | (u >> ((64 - n) & 63))
        else:
            r = (u >> n) | (u << ((64 - n) & 63))
        return _signed64(r)
    if f == "real":
        return to_double(a0)
    if f == "imag":
        return 0.0
    raise _EvalError("Unsupported or unknown function")


def _eval(node):
    kind = node[0]
    if kind == "int":
        try:
            return java_parse_int(node[1], 64)
        except JvmNumberFormatError as e:
            raise _EvalError(str(e)) from None
    if kind == "float":
        return java_parse_double(node[1])
    if kind == "bool":
        return node[1]
    if kind == "paren":
        return _eval(node[1])
    if kind == "ident":
        if node[1] in CONSTANTS:
            return CONSTANTS[node[1]]
        raise _EvalError("Unknown identifier in constant expression")
    if kind == "hwq":
        raise _EvalError("Hardware qubit not valid in const expression")
    if kind == "lit":
        raise _EvalError(f"Cannot evaluate {node[1]} literal {node[2]}")
    if kind == "call":
        return apply_function(node[1], [_eval(a) for a in node[2]])
    if kind == "unary":
        op, v = node[1], _eval(node[2])
        if op == "-":
            return -to_double(v)
        if op == "+":
            return to_double(v)
        if op == "!":
            return not to_bool(v)
        return _signed64(~to_long(v))
    if kind == "pow":
        return java_pow(to_double(_eval(node[1])), to_double(_eval(node[2])))
    if kind == "bin":
        op, a, b = node[1], _eval(node[2]), _eval(node[3])
        return _binary(op, a, b)
    raise _EvalError(f"Cannot evaluate {kind}")


def _binary(op, a, b):
    if op in ("*", "+", "-"):
        if _is_double(a) or _is_double(b):
            x, y = to_double(a), to_double(b)
            return x * y if op == "*" else x + y if op == "+" else x - y
        x, y = to_long(a), to_long(b)
        return _long(x * y if op == "*" else x + y if op == "+" else x - y)
    if op == "/":
        return java_div(to_double(a), to_double(b))
    if op == "%":
        if isinstance(a, int) and not isinstance(a, bool) and isinstance(b, int) and not isinstance(b, bool):
            return _clj_mod(to_long(a), to_long(b))
        return _double_mod(a, b)
    if op in ("<", ">", "<=", ">="):
        x, y = to_double(a), to_double(b)
        return {"<": x < y, ">": x > y, "<=": x <= y, ">=": x >= y}[op]
    if op == "==":
        return to_double(a) == to_double(b)
    if op == "!=":
        return to_double(a) != to_double(b)
    if op == "&":
        return to_long(a) & to_long(b)
    if op == "^":
        return to_long(a) ^ to_long(b)
    if op == "|":
        return to_long(a) | to_long(b)
    if op == "&&":
        return to_bool(a) and to_bool(b)
    if op == "||":
        return to_bool(a) or to_bool(b)
    raise _EvalError("Unknown operator " + op)


def evaluate_expression(node):
    """evaluate-expression (qasm3_parser.clj:341-512): the value (int, float
    or bool) of a parsed constant expression."""
    try:
        return _eval(node)
    except (_EvalError, JvmNumberFormatError, ZeroDivisionError, OverflowError, RecursionError) as e:
        raise QasmError("Failed to evaluate OpenQASM expression", {"expr": node, "cause": str(e)}) from None


def parse_expression(text: str):
    """Parse one expression (the grammar's ``expression``) to its node."""
    p = _Parser(text)
    try:
        node = p.expression()
    except _Fail:
        raise QasmParseError("Parse error", {"text": text}) from None
    if not p.at_end():
        raise QasmParseError("Parse error", {"text": text})
    return node


def _find_int_node(node):
    if isinstance(node, tuple) and node:
        if node[0] == "int":
            return node
        children = node[1:]
    elif isinstance(node, list):
        children = node
    else:
        return None
    for child in children:
        found = _find_int_node(child)
        if found is not None:
            return found
    return None


def _first_int_literal(node):
    """parse-simple-integer (qasm3_parser.clj:185-196): the value of the first
    decimal integer literal of a subtree in source order; None if there is
    none or it does not fit a 32-bit int."""
    found = _find_int_node(node)
    if found is None:
        return None
    try:
        return java_parse_int(found[1])
    except JvmNumberFormatError:
        return None


def _operand_index(operand):
    """extract-qubit-index (qasm3_parser.clj:524-534)."""
    if operand[0] != "indexed" or not operand[2]:
        return None
    return _first_int_literal(operand[2][0])


# ---- gate calls -> operations (qasm3_parser.clj:564-674) ------------------------------------------
_ONE = {"h", "x", "y", "z", "s", "s-dag", "t", "t-dag"}
_ROT = {"rx", "ry", "rz", "phase"}
_TWO = {"cnot": ("control", "target"), "cz": ("control", "target"), "cy": ("control", "target"),
        "swap": ("qubit1", "qubit2"), "iswap": ("qubit1", "qubit2")}
_CROT = {"crx", "cry", "crz"}


def _param0(params):
    p = params[0] if params else None
    return p if p is not None and p is not False else 0.0


def parse_gate_call(name, params: list, operands: list):
    """parse-gate-call: the operation map for one gate call, or None for
    calls with modifiers, unknown names or too few operands."""
    if not name:
        log.warning("Unknown gate '%s' not found in operation registry", name)
        return None
    canonical = qc.resolve_gate_alias(name)
    if canonical not in qc.OPERATION_CATALOG:
        log.warning("Unknown gate '%s' not found in operation registry", name)
        return None
    ops = [i for i in (_operand_index(o) for o in operands) if i is not None]

    def at(k):
        if k >= len(ops):
            raise IndexError(k)
        return ops[k]

    first = ops[0] if ops else None
    second = ops[1] if len(ops) > 1 else None
    try:
        if canonical in _ONE:
            return {"operation-type": canonical, "operation-params": {"target": first}}
        if canonical in _ROT:
            return {"operation-type": canonical, "operation-params": {"target": first, "angle": _param0(params)}}
        if canonical in _TWO:
            k1, k2 = _TWO[canonical]
            return {"operation-type": canonical, "operation-params": {k1: first, k2: second}}
        if canonical in _CROT:
            return {"operation-type": canonical,
                    "operation-params": {"control": first, "target": second, "angle": _param0(params)}}
        if canonical == "toffoli":
            return {"operation-type": "toffoli",
                    "operation-params": {"control1": first, "control2": second, "target": at(2)}}
        if canonical == "fredkin":
            return {"operation-type": "fredkin",
                    "operation-params": {"control": first, "target1": second, "target2": at(2)}}
        if canonical == "rydberg-cz":
            return {"operation-type": canonical, "operation-params": {"control": first, "target": second}}
        if canonical == "rydberg-cphase":
            return {"operation-type": canonical,
                    "operation-params": {"control": first, "target": second, "angle": _param0(params)}}
        if canonical.startswith("global-"):
            p = {"angle": _param0(params)} if canonical in ("global-rx", "global-ry", "global-rz") else {}
            return {"operation-type": canonical, **({"operation-params": p} if p else {})}
        log.info("Using generic add-gate for: :%s", canonical)
        if len(ops) == 1:
            p = {"target": ops[0]}
        elif len(ops) == 2:
            p = {"control": ops[0], "target": ops[1]}
        elif len(ops) == 3:
            p = {"control1": ops[0], "control2": ops[1], "target": ops[2]}
        else:
            p = {}
        return {"operation-type": canonical, **({"operation-params": p} if p else {})}
    except IndexError:
        log.warning("Error creating gate :%s: index out of range", canonical)
        return None


def qasm_to_circuit(qasm_code: str) -> dict:
    """qasm-to-circuit (qasm3_parser.clj:741-758) with transform-parse-tree
    (qasm3_parser.clj:676-739)."""
    raw_specs = collect_result_specs_from_qasm(split_lines(qasm_code))
    pre = preprocess_qasm(qasm_code)
    p = _Parser(pre)
    try:
        p.program()
    except _Fail:
        raise QasmParseError("Parse error", {"position": p.i}) from None
    except RecursionError:
        raise QasmParseError("Parse error", {"position": p.i}) from None
    num_qubits = 0
    for _, designator in p.decls:
        if designator is None:
            raise QasmError("Wrong number of args (3) passed to the quantum declaration transform")
        size = _first_int_literal(designator)
        if size is None:
            raise QasmError("qubit register size must contain an integer literal", {"designator": designator})
        num_qubits = size
        break
    operations = []
    for _, name, params, operands in p.calls:
        values = [evaluate_expression(e) for e in params]
        op = parse_gate_call(name, values, operands)
        if op is not None:
            operations.append(op)
    if operations and num_qubits <= 0:
        raise QasmError("Assert failed: (s/valid? ::circuit circuit) — no qubit register declared")
    for k, op in enumerate(operations[:-1]):
        a = (op.get("operation-params") or {}).get("angle")
        if a is not None and (isinstance(a, bool) or not isinstance(a, (int, float))):
            raise QasmError("Assert failed: (s/valid? ::circuit circuit) — angle is not a number")
    c = qc.create_circuit(num_qubits, "Parsed QASM3 Circuit")
    c["operations"] = operations
    specs = collect_result_specs_from_qasm(split_lines(pre))
    specs.update(raw_specs)
    c["result-specs"] = specs
    return c


# ---- emission (qasm3_parser.clj:66-154, 760-857) ------------------------------------------------------
GATE_INFO = {
    # operation-type, operation-name (operation_registry.clj:55-363)
    "x": ("single-qubit", "X"), "y": ("single-qubit", "Y"), "z": ("single-qubit", "Z"),
    "i": ("single-qubit", "I"), "h": ("single-qubit", "H"), "s": ("single-qubit", "S"),
    "s-dag": ("single-qubit", "S†"), "t": ("single-qubit", "T"), "t-dag": ("single-qubit", "T†"),
    "rx": ("parametric", "RX"), "ry": ("parametric", "RY"), "rz": ("parametric", "RZ"),
    "phase": ("parametric", "Phase"), "cnot": ("two-qubit", "CNOT"), "cz": ("two-qubit", "CZ"),
    "cy": ("two-qubit", "CY"), "swap": ("two-qubit", "SWAP"), "iswap": ("two-qubit", "iSWAP"),
    "crx": ("parametric", "CRX"), "cry": ("parametric", "CRY"), "crz": ("parametric", "CRZ"),
    "toffoli": ("multi-qubit", "Toffoli"), "fredkin": ("multi-qubit", "Fredkin"),
    "rydberg-cz": ("two-qubit", "Rydberg CZ"), "rydberg-cphase": ("parametric", "Rydberg CPhase"),
    "rydberg-blockade": ("multi-qubit", "Rydberg Blockade"),
    "global-rx": ("global-parametric", "Global RX"), "global-ry": ("global-parametric", "Global RY"),
    "global-rz": ("global-parametric", "Global RZ"), "global-h": ("global", "Global H"),
    "global-x": ("global", "Global X"), "global-y": ("global", "Global Y"), "global-z": ("global", "Global Z"),
    "measure": ("single-qubit", "Measure"),
}


def emit_qasm_pragmas(options) -> str | None:
    """emit-qasm-pragmas (qasm3_parser.clj:66-154): reads ``result-specs``
    and ``target`` from its argument; singular ``observable``/``target`` are
    accepted; lines without a comment header."""
    options = options or {}
    result_specs = options.get("result-specs") if isinstance(options, dict) else None
    target = options.get("target", "qclojure")
    if not result_specs:
        return None
    head = "#pragma " + clj_str(PRAGMA_TARGET.get(target))

    def plural(spec, many, one):
        v = spec.get(many)
        return list(v) if v is not None and v is not False else [spec.get(one)]

    lines: list[str] = []
    for rtype, spec in result_specs.items():
        spec = spec if isinstance(spec, dict) else {}
        if rtype == "measurement":
            s = head + " result measurement shots=" + clj_str(spec.get("shots", 1000))
            q = spec.get("qubits")
            if q is not None and q is not False:
                s += " qubits=" + ",".join(clj_str(x) for x in q)
            lines.append(s)
        elif rtype in ("expectation", "variance", "sample"):
            obs, tg = plural(spec, "observables", "observable"), plural(spec, "targets", "target")
            for i, o in enumerate(obs):
                if o is None:
                    raise QasmError("NullPointerException: observable is nil")
                s = f"{head} result {rtype} observable=" + clj_name(o)
                if rtype == "sample":
                    s += " shots=" + clj_str(spec.get("shots", 1000))
                if i < len(tg):
                    s += " target=" + clj_str(tg[i])
                lines.append(s)
        elif rtype == "probability":
            s = head + " result probability"
            t = spec.get("targets")
            if t is not None and t is not False:
                s += " targets=" + ",".join(clj_str(x) for x in t)
            st = spec.get("states")
            if st is not None and st is not False:
                s += " states=" + ",".join(clj_str(x) for x in st)
            lines.append(s)
        elif rtype == "amplitude":
            lines.append(head + " result amplitude states=" + ",".join(clj_str(x) for x in spec.get("states", []) or []))
        elif rtype == "state-vector":
            lines.append(head + " result state_vector // Simulation-only result")
        elif rtype == "density-matrix":
            lines.append(head + " result density_matrix // Simulation-only result")
        elif rtype == "fidelity":
            lines.append(head + " result fidelity // Simulation-only result")
        else:
            lines.append("// Unknown result type: " + clj_name(rtype))
    if not lines:
        return None
    return "\n" + "\n".join(lines) + "\n"


def gate_to_qasm3(gate: dict, options: dict | None = None) -> str:
    """gate-to-qasm3 (qasm3_parser.clj:760-824): names from the operation
    registry (lower-cased operation-name, with p/sdg/tdg), shapes by its
    operation-type."""
    gtype = gate.get("operation-type")
    p = gate.get("operation-params") or {}
    info = GATE_INFO.get(qc.resolve_gate_alias(gtype)) if isinstance(gtype, str) else None
    if info is None:
        return "// Unknown gate: " + clj_name(gtype)
    otype, oname = info
    qname = {"phase": "p", "s-dag": "sdg", "t-dag": "tdg"}.get(gtype, oname.lower())
    q = lambda v: "q[" + clj_str(v) + "]"  # noqa: E731
    if otype == "single-qubit":
        return f"{qname} {q(p.get('target'))};"
    if otype == "parametric":
        return f"{qname}({clj_str(p.get('angle'))}) {q(p.get('target'))};"
    if otype == "two-qubit":
        if gtype == "cnot":
            return f"cx {q(p.get('control'))}, {q(p.get('target'))};"
        if gtype in ("swap", "iswap"):
            return f"{gtype} {q(p.get('qubit1'))}, {q(p.get('qubit2'))};"
        return f"{qname} {q(p.get('control'))}, {q(p.get('target'))};"
    if otype == "multi-qubit":
        if gtype == "toffoli":
            return f"ccx {q(p.get('control1'))}, {q(p.get('control2'))}, {q(p.get('target'))};"
        if gtype == "fredkin":
            return f"cswap {q(p.get('control'))}, {q(p.get('target1'))}, {q(p.get('target2'))};"
        return "// Multi-qubit gate: " + clj_name(gtype)
    if otype == "global":
        return "// Global gate: " + clj_name(gtype)
    return f"// Global parametric gate: {clj_name(gtype)}({clj_str(p.get('angle'))})"


def circuit_to_qasm(circuit: dict, options: dict | None = None) -> str:
    """circuit-to-qasm (qasm3_parser.clj:826-857): header, one line per
    operation, then the pragmas of ``options["result-specs"]`` (which, as in
    the reference, are read from that map's own "result-specs" key)."""
    options = options or {}
    target = options.get("target", "qclojure")
    n = clj_str(circuit.get("num-qubits"))
    header = "OPENQASM 3.0;\n" + ("" if target == "braket" else 'include "stdgates.inc";\n\n') + \
        f"qubit[{n}] q;\n" + f"bit[{n}] c;\n\n"
    gates = "\n".join(gate_to_qasm3(op, options) for op in circuit.get("operations") or [])
    rs = options.get("result-specs")
    pragmas = emit_qasm_pragmas(rs) if rs is not None and rs is not False else None
    return header + gates + (pragmas or "")
