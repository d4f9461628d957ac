"""awkward_amd -- an MI355X-native ``hip`` backend for Awkward Array's
jagged-array kernel layer.
"""
