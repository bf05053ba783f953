// state.hpp -- the game types of gpu_ai's interface (src/state.hpp:10-312), restated over
// 32-bit bitboards so that a State converts to the engine's 16-byte mi_pos without copying a
// 776-byte board.  The members and methods the reference's callers use are kept with their names
// and meaning: turn, movesSinceLastCapture, operator[], ==, !=, <, getNextTurn, move, genMoves,
// getMoves, isGameOver, getWinner (src/mcts.cpp, src/player.cpp, src/driver.cpp).
#pragma once

#include <stdint.h>

#include <iosfwd>
#include <vector>

#include "gpuai/mi_playout.h"

#define BOARD_SIZE 8
#define NUM_LOCS 32
#define NUM_PLAYERS 2
#define NUM_DRAW_MOVES 50
#define MAX_MOVE_JUMPS 8
#define MAX_MOVES 100

enum PlayerId { PLAYER_1, PLAYER_2, PLAYER_NONE = -1 };

enum PieceType { CHECKER, CHECKER_KING };

// A board location (src/state.hpp:34-68).
struct Loc {
  uint8_t row = 0, col = 0;
  Loc() = default;
  Loc(uint8_t row, uint8_t col) : row(row), col(col) {}
  bool operator==(const Loc &o) const { return row == o.row && col == o.col; }
  bool operator!=(const Loc &o) const { return !(*this == o); }
  bool isValid() const { return row < BOARD_SIZE && col < BOARD_SIZE; }
  // Dark-square index s = 4*row + col/2 (src/state.cu:173-176).
  uint8_t square() const { return (uint8_t)(4 * row + (col >> 1)); }
  static Loc ofSquare(unsigned s) {
    uint8_t row = (uint8_t)(s >> 2);
    return Loc(row, (uint8_t)(2 * (s & 3) + ((row & 1) == 0 ? 1 : 0)));
  }
};

// Contents of a square (src/state.hpp:71-108).
struct BoardItem {
  bool occupied;
  PieceType type;
  PlayerId owner;
  bool operator==(const BoardItem &o) const {
    return occupied == o.occupied && type == o.type && owner == o.owner;
  }
  bool operator!=(const BoardItem &o) const { return !(*this == o); }
};

struct Move;

struct State {
  uint32_t p1 = 0, p2 = 0, kings = 0;  // bitboards over the 32 dark squares
  PlayerId turn = PLAYER_1;
  unsigned movesSinceLastCapture = 0;

  BoardItem operator[](Loc loc) const;
  bool operator==(const State &o) const {
    return p1 == o.p1 && p2 == o.p2 && kings == o.kings && turn == o.turn &&
           movesSinceLastCapture == o.movesSinceLastCapture;
  }
  bool operator!=(const State &o) const { return !(*this == o); }
  bool operator<(const State &o) const;

  PlayerId getNextTurn() const;
  // Applies a move (src/state.cu:57-92).
  void move(const Move &move);
  // The legal moves in the reference's order (src/state.cu:239-245); returns their number.
  uint8_t genMoves(Move result[MAX_MOVES]) const;
  std::vector<Move> getMoves() const;
  bool isGameOver() const;    // src/state.cpp:16-18
  PlayerId getWinner() const; // src/state.cpp:20-23

  mi_pos pack() const;
  static State unpack(const mi_pos &p);
};

// A move (src/state.hpp:253-296).  Equality compares the first `jumps` entries only
// (src/state.cu:440-454).
struct Move {
  Loc from, to;
  Loc removed[MAX_MOVE_JUMPS];
  Loc intermediate[MAX_MOVE_JUMPS];
  uint8_t jumps = 0;
  bool promoted = false;

  Move() = default;
  Move(Loc from, Loc to, uint8_t jumps = 0, bool promoted = false)
      : from(from), to(to), jumps(jumps), promoted(promoted) {}
  bool operator==(const Move &o) const;
  bool operator!=(const Move &o) const { return !(*this == o); }
  // Appends a jump landing on newTo (src/state.cu:456-462).
  void addJump(Loc newTo);
};

// The initial position (src/state.cpp:25-40).
State getStartingState();

std::ostream &operator<<(std::ostream &, PlayerId);
std::ostream &operator<<(std::ostream &, Loc);
std::ostream &operator<<(std::ostream &, const Move &);
std::ostream &operator<<(std::ostream &, const State &);
