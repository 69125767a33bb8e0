"""Lexical helpers for C/C++ OP2 application sources (no compiler front end):
comment stripping that keeps offsets, balanced call extraction, top-level
argument splitting and function-definition lookup."""
import re


def strip_comments(text):
    """Replace comments by spaces (newlines kept) so offsets still match."""
    out = list(text)
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c in "\"'":
            j = i + 1
            while j < n and text[j] != c:
                j += 2 if text[j] == "\\" else 1
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            j = n if j < 0 else j
            for k in range(i, j):
                out[k] = " "
            i = j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
            for k in range(i, j):
                if out[k] != "\n":
                    out[k] = " "
            i = j
        else:
            i += 1
    return "".join(out)


def match_close(text, open_pos):
    """Offset of the bracket closing the one at open_pos (strings skipped)."""
    pairs = {"(": ")", "[": "]", "{": "}"}
    stack = [pairs[text[open_pos]]]
    i = open_pos + 1
    while i < len(text):
        c = text[i]
        if c in "\"'":
            j = i + 1
            while j < len(text) and text[j] != c:
                j += 2 if text[j] == "\\" else 1
            i = j + 1
            continue
        if c in pairs:
            stack.append(pairs[c])
        elif c in ")]}":
            if c != stack.pop():
                raise SyntaxError(f"unbalanced '{c}' at offset {i}")
            if not stack:
                return i
        i += 1
    raise SyntaxError(f"unclosed '{text[open_pos]}' at offset {open_pos}")


def split_args(s):
    """Split an argument list at top-level commas."""
    args, depth, cur, i = [], 0, [], 0
    while i < len(s):
        c = s[i]
        if c in "\"'":
            j = i + 1
            while j < len(s) and s[j] != c:
                j += 2 if s[j] == "\\" else 1
            cur.append(s[i:j + 1])
            i = j + 1
            continue
        if c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
        if c == "," and depth == 0:
            args.append("".join(cur).strip())
            cur = []
        else:
            cur.append(c)
        i += 1
    last = "".join(cur).strip()
    if last or args:
        args.append(last)
    return args


def find_calls(text, name):
    """[(start, end, args)] for every call `name(...)`; end is one past ')'."""
    out = []
    for m in re.finditer(r"\b%s\s*\(" % re.escape(name), text):
        if m.start() > 0 and (text[m.start() - 1] == "." or text[m.start() - 2:m.start()] == "->"):
            continue
        open_pos = m.end() - 1
        close = match_close(text, open_pos)
        out.append((m.start(), close + 1, split_args(text[open_pos + 1:close])))
    return out


def string_literal(s):
    s = s.strip()
    if len(s) < 2 or s[0] != '"' or s[-1] != '"':
        raise ValueError(f"expected a string literal, got {s!r}")
    return s[1:-1]


def find_function(text, name):
    """(start, end) of the definition of function `name` (declaration
    specifiers through the closing brace), or None."""
    clean = strip_comments(text)
    for m in re.finditer(r"\b%s\s*\(" % re.escape(name), clean):
        close = match_close(clean, m.end() - 1)
        k = close + 1
        while k < len(clean) and clean[k] in " \t\r\n":
            k += 1
        if k >= len(clean) or clean[k] != "{":
            continue  # a call or a prototype
        # the return type starts after the previous ';', '}' or preprocessor line
        start = m.start()
        while start > 0 and clean[start - 1] not in ";}\n":
            start -= 1
        while True:
            prev = clean.rfind("\n", 0, max(start - 1, 0))
            line = clean[prev + 1:start].strip()
            if start > 0 and line and not line.startswith("#") and line[-1] not in ";}":
                start = prev + 1
            else:
                break
        return start, match_close(clean, k) + 1
    return None


def int_constants(text):
    """Integer macros and constants (`#define N 4`, `const int N = 4;`)."""
    out = {}
    for m in re.finditer(r"^\s*#\s*define\s+([A-Za-z_]\w*)\s+\(?\s*(-?\d+)\s*\)?\s*$", text, re.M):
        out[m.group(1)] = int(m.group(2))
    for m in re.finditer(r"\b(?:static\s+)?(?:const|constexpr)\s+(?:int|long|size_t|unsigned)\s+"
                         r"([A-Za-z_]\w*)\s*=\s*(-?\d+)\s*;", text):
        out[m.group(1)] = int(m.group(2))
    return out
