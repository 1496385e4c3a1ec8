"""The index scan's carried state, byte by byte (test helper): escape parity
(a backslash that is not itself escaped escapes the next byte), in-string
parity (an unescaped quote toggles it), nesting depth as int8 (brackets
outside strings, escaped or not; the document starts at -1) and the newline
count.  This is
the definition the string mask and the leveled rows are built from
(string-index.cu:3-31, leveled-bitmaps-index.cu:28-71), restated as one
sequential scan for small inputs."""


def scan(data: bytes, esc: int = 0, instr: int = 0, depth: int = -1, nl: int = 0):
    """The state after `data` from the given one."""
    for b in data:
        if b == 0x0A:
            nl += 1
        escaped, esc = esc, 0
        if b == 0x5C and not escaped:  # a backslash escapes the next byte
            esc = 1
        elif b == 0x22 and not escaped:  # an unescaped quote
            instr ^= 1
        elif not instr and b in (0x7B, 0x5B):  # brackets count escaped or not
            depth = (depth + 1 + 128) % 256 - 128
        elif not instr and b in (0x7D, 0x5D):
            depth = (depth - 1 + 128) % 256 - 128
    return esc, instr, depth, nl


def summary(data: bytes):
    """(table, newlines) as gpj_shard_summary encodes them."""
    table = 0
    nl = 0
    for s in range(4):
        e, st, d, nl = scan(data, s & 1, s >> 1, 0, 0)
        table |= (e | (st << 1)) << (8 * s)
        table |= (d % 256) << (32 + 8 * s)
    return table, nl
