// state.cpp -- game types over bitboards; move generation through the same lane logic the
// device kernels run (gpu_ai_amd/csrc/checkers.h), here executed on the host for the callers
// (tree expansion in MCTS, the CLI), never for playouts.
#include "state.hpp"

#include <cstring>
#include <ostream>

#include "checkers.h"

BoardItem State::operator[](Loc loc) const {
  BoardItem item = {false, CHECKER, PLAYER_1};
  if (!loc.isValid() || ((loc.row + loc.col) & 1) == 0) return item;
  const uint32_t bit = 1u << loc.square();
  if ((p1 | p2) & bit) {
    item.occupied = true;
    item.owner = (p1 & bit) ? PLAYER_1 : PLAYER_2;
    item.type = (kings & bit) ? CHECKER_KING : CHECKER;
  }
  return item;
}

// Square-by-square comparison in row-major order, then turn, then the draw counter, with
// BoardItem ordered by (occupied, type, owner) as src/state.cu:13-44.
bool State::operator<(const State &o) const {
  for (unsigned s = 0; s < NUM_LOCS; s++) {
    const BoardItem a = (*this)[Loc::ofSquare(s)], b = o[Loc::ofSquare(s)];
    if (a == b) continue;
    if (a.occupied != b.occupied) return a.occupied < b.occupied;
    if (a.type != b.type) return a.type < b.type;
    return a.owner < b.owner;
  }
  if (turn != o.turn) return turn < o.turn;
  return movesSinceLastCapture < o.movesSinceLastCapture;
}

PlayerId State::getNextTurn() const {
  switch (turn) {
    case PLAYER_1: return PLAYER_2;
    case PLAYER_2: return PLAYER_1;
    default: return PLAYER_NONE;
  }
}

void State::move(const Move &m) {
  const uint32_t fb = 1u << m.from.square(), tb = 1u << m.to.square();
  uint32_t removed = 0;
  for (unsigned i = 0; i < m.jumps; i++) removed |= 1u << m.removed[i].square();
  uint32_t &own = turn == PLAYER_1 ? p1 : p2;
  uint32_t &opp = turn == PLAYER_1 ? p2 : p1;
  const bool wasKing = (kings & fb) != 0;
  opp &= ~removed;
  kings &= ~removed;
  own = (own & ~fb) | tb;
  kings &= ~fb;
  if (wasKing || m.promoted) kings |= tb;
  turn = getNextTurn();
  movesSinceLastCapture = m.jumps > 0 ? 0 : movesSinceLastCapture + 1;
}

uint8_t State::genMoves(Move result[MAX_MOVES]) const {
  const mi_pos p = pack();
  const ck::Lane L = ck::lane_from_abs(p.p1, p.p2, p.kings, p.turn, 0u);
  ck::Gen g;
  ck::analyse(L, g);
  const uint32_t n = g.capture ? ck::count_walks(L, g) : ck::slot_count(g.slot);
  const uint32_t keep = n < MAX_MOVES ? n : MAX_MOVES;
  auto visit = [&](uint32_t idx, uint32_t from, uint32_t to, uint32_t, uint32_t jumps,
                   uint32_t dirs, uint32_t king) {
    const uint32_t ref = L.turn ? n - 1u - idx : idx;
    if (ref >= keep) return;
    Move m(Loc::ofSquare(ck::abs_sq(L, from)), Loc::ofSquare(ck::abs_sq(L, from)));
    uint32_t x = from;
    for (uint32_t t = 0; t < jumps; t++) {
      const uint32_t land = (uint32_t)((int)x + ck::jump_off((dirs >> (2u * t)) & 3u));
      m.addJump(Loc::ofSquare(ck::abs_sq(L, land)));
      x = land;
    }
    m.to = Loc::ofSquare(ck::abs_sq(L, to));
    m.promoted = !king && to >= ck::FAR_ROW_FIRST;
    result[ref] = m;
  };
  ck::enumerate(L, g, visit);
  return (uint8_t)keep;
}

std::vector<Move> State::getMoves() const {
  Move result[MAX_MOVES];
  const uint8_t n = genMoves(result);
  return std::vector<Move>(result, result + n);
}

bool State::isGameOver() const {
  if (movesSinceLastCapture >= NUM_DRAW_MOVES) return true;
  const mi_pos p = pack();
  return ck::game_over(ck::lane_from_abs(p.p1, p.p2, p.kings, p.turn, 0u));
}

PlayerId State::getWinner() const {
  return movesSinceLastCapture < NUM_DRAW_MOVES ? getNextTurn() : PLAYER_NONE;
}

mi_pos State::pack() const {
  mi_pos p;
  std::memset(&p, 0, sizeof(p));
  p.p1 = p1;
  p.p2 = p2;
  p.kings = kings & (p1 | p2);
  p.turn = turn == PLAYER_2 ? 1 : 0;
  p.counter = (uint8_t)(movesSinceLastCapture > 255 ? 255 : movesSinceLastCapture);
  return p;
}

State State::unpack(const mi_pos &p) {
  State s;
  s.p1 = p.p1;
  s.p2 = p.p2;
  s.kings = p.kings & (p.p1 | p.p2);
  s.turn = p.turn ? PLAYER_2 : PLAYER_1;
  s.movesSinceLastCapture = p.counter;
  return s;
}

bool Move::operator==(const Move &o) const {
  if (from != o.from || to != o.to || jumps != o.jumps || promoted != o.promoted) return false;
  for (unsigned i = 0; i < jumps; i++)
    if (removed[i] != o.removed[i] || intermediate[i] != o.intermediate[i]) return false;
  return true;
}

void Move::addJump(Loc newTo) {
  if (jumps >= MAX_MOVE_JUMPS) return;
  intermediate[jumps] = newTo;
  removed[jumps] = Loc((uint8_t)((newTo.row - to.row) / 2 + to.row),
                       (uint8_t)((newTo.col - to.col) / 2 + to.col));
  jumps++;
  to = newTo;
}

State getStartingState() {
  State s;
  s.p1 = 0x00000FFFu;
  s.p2 = 0xFFF00000u;
  s.turn = PLAYER_1;
  s.movesSinceLastCapture = 0;
  return s;
}

std::ostream &operator<<(std::ostream &os, PlayerId p) {
  switch (p) {
    case PLAYER_1: return os << "Player 1";
    case PLAYER_2: return os << "Player 2";
    default: return os << "Player NONE";
  }
}

// Files a-h left to right, ranks 8-1 top to bottom (src/state.cpp:66-68).
std::ostream &operator<<(std::ostream &os, Loc loc) {
  return os << (char)('a' + loc.col) << (char)('0' + (BOARD_SIZE - loc.row));
}

std::ostream &operator<<(std::ostream &os, const Move &m) {
  os << m.from;
  for (unsigned i = 0; i + 1 < m.jumps; i++) os << " " << m.intermediate[i];
  os << " " << m.to;
  if (m.promoted) os << " (king)";
  return os;
}

std::ostream &operator<<(std::ostream &os, const State &s) {
  for (int row = 0; row < BOARD_SIZE; row++) {
    os << (char)('0' + (BOARD_SIZE - row)) << " ";
    for (int col = 0; col < BOARD_SIZE; col++) {
      const BoardItem it = s[Loc((uint8_t)row, (uint8_t)col)];
      char c = '.';
      if (it.occupied) c = it.owner == PLAYER_1 ? (it.type == CHECKER_KING ? 'X' : 'x')
                                                : (it.type == CHECKER_KING ? 'O' : 'o');
      else if (((row + col) & 1) == 0) c = ' ';
      os << c << ' ';
    }
    os << "\n";
  }
  os << "  a b c d e f g h\n" << s.turn << " to move, " << s.movesSinceLastCapture
     << " moves since the last capture\n";
  return os;
}
