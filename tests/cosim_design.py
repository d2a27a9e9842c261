"""A small SPI master in AIGPDK gate-level Verilog, for the co-simulation
tests: a 7-bit saturating counter drives a single-bit SPI read of two bytes
at a fixed address (command 0x03, 24-bit address, MSB first), and a 16-bit
shift register captures the flash's MISO bits.  Everything is a function of
the counter, generated as multiplexer trees from truth tables.

Timeline in counter values c (one per system clock cycle):
  c in [2, 98)      chip select low
  c odd in window   SPI clock high
  c in [2, 66)      MOSI = bit 31 - (c - 2) / 2 of (0x03 << 24 | address)
  c odd, 67..97     the MISO input shifts into rx (16 samples)
After the transaction rx (rx15..rx0) holds image[address] << 8 |
image[address + 1].
"""

CS_LO, CS_HI = 2, 98
CMD_END = 66
RX_FIRST, RX_LAST = 67, 97


class _Gates:
    def __init__(self):
        self.lines = []
        self.wires = []
        self.n = 0

    def _new(self):
        self.n += 1
        name = f"n{self.n}"
        self.wires.append(name)
        return name

    def and2(self, a, b, ia=0, ib=0, iy=0):
        y = self._new()
        self.lines.append(f"  AND2_{ia}{ib}_{iy} g{self.n} (.A({a}), .B({b}), .Y({y}));")
        return y

    def inv(self, a):
        y = self._new()
        self.lines.append(f"  INV g{self.n} (.A({a}), .Y({y}));")
        return y

    def mux(self, s, f1, f0):
        """s ? f1 : f0 with constants (Python bools) folded."""
        if isinstance(f1, bool) and isinstance(f0, bool):
            if f1 == f0:
                return f1
            return s if f1 else self.inv(s)
        if f1 is True:
            return self.and2(s, f0, 1, 1, 1)          # s | f0
        if f1 is False:
            return self.and2(s, f0, 1, 0, 0)          # ~s & f0
        if f0 is True:
            return self.and2(s, f1, 0, 1, 1)          # ~s | f1
        if f0 is False:
            return self.and2(s, f1)                   # s & f1
        if f1 == f0:
            return f0
        a = self.and2(s, f1)
        b = self.and2(s, f0, 1, 0, 0)
        return self.and2(a, b, 1, 1, 1)               # a | b

    def func(self, table, bits, memo):
        """Truth table over bits (bits[0] least significant) as gates."""
        key = (tuple(table), len(bits))
        if key in memo:
            return memo[key]
        if all(table) or not any(table):
            r = bool(table[0])
        else:
            half = len(table) // 2
            f0 = self.func(table[:half], bits[:-1], memo)
            f1 = self.func(table[half:], bits[:-1], memo)
            r = self.mux(bits[-1], f1, f0)
        memo[key] = r
        return r


def spi_master_verilog(address):
    word = (0x03 << 24) | (address & 0xFFFFFF)
    g = _Gates()
    cbits = [f"c{i}" for i in range(7)]
    rx = [f"rx{i}" for i in range(16)]
    memo = {}

    def f(pred):
        return g.func([bool(pred(c)) for c in range(128)], cbits, memo)

    csn = f(lambda c: not (CS_LO <= c < CS_HI))
    sclk = f(lambda c: CS_LO <= c < CS_HI and c & 1)
    mosi = f(lambda c: CS_LO <= c < CMD_END and (word >> (31 - (c - CS_LO) // 2)) & 1)
    shift = f(lambda c: RX_FIRST <= c <= RX_LAST and c & 1)
    dnext = []
    for i in range(7):
        nxt = f(lambda c, i=i: ((min(c + 1, 127)) >> i) & 1)
        dnext.append(nxt if isinstance(nxt, str) else None)
        if dnext[-1] is None:
            raise AssertionError("constant counter bit")
        dnext[-1] = g.and2(nxt, "rst", 0, 1, 0)  # held at 0 in reset
    rxn = [g.mux(shift, "miso" if i == 0 else rx[i - 1], rx[i]) for i in range(16)]

    def net(x):
        if isinstance(x, bool):
            return "1'b1" if x else "1'b0"
        return x

    outs = ["flash_clk", "flash_csn", "flash_d0"] + [f"rx_o{i}" for i in range(16)]
    ports = ["clk", "rst", "miso"] + outs
    L = [f"module spi_master({', '.join(ports)});"]
    L += ["  input clk;", "  input rst;", "  input miso;"]
    L += [f"  output {o};" for o in outs]
    wires = cbits + rx + [f"d{i}" for i in range(7)] + [f"rxd{i}" for i in range(16)] + g.wires
    L += [f"  wire {w};" for w in wires]
    L += g.lines
    for i in range(7):
        L.append(f"  assign d{i} = {dnext[i]};")
        L.append(f"  DFF q{i} (.CLK(clk), .D(d{i}), .Q(c{i}));")
    for i in range(16):
        L.append(f"  assign rxd{i} = {net(rxn[i])};")
        L.append(f"  DFF r{i} (.CLK(clk), .D(rxd{i}), .Q(rx{i}));")
    L.append(f"  assign flash_clk = {net(sclk)};")
    L.append(f"  assign flash_csn = {net(csn)};")
    L.append(f"  assign flash_d0 = {net(mosi)};")
    for i in range(16):
        L.append(f"  assign rx_o{i} = rx{i};")
    L.append("endmodule")
    return "\n".join(L) + "\n"
